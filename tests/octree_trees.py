"""Octree test trees built with the oracle's refinement (shared by the CPU and
GPU octree multigrid tests)."""
import numpy as np


def adaptive_octree(orc):
    """Anisotropic box (3 x 2 x 1.5), base level 4, a box refined twice, a
    seeded refinement storm up to level 7 (balanced by refineByMorton) and an
    InsideSolid block: about 31k leaves, multigrid levels
    31032 -> 30829 -> 8926 -> 4096 -> 512."""
    t = orc.octree((-1.0, -2.0, 0.5), (3.0, 2.0, 1.5), 7)
    t.refine_uniform(4)
    t.refine_box((-0.6, -1.6, 0.6), (0.6, -0.6, 1.3), 2)
    rng = np.random.default_rng(777)
    for _ in range(40):
        k, l, _m = t.dump()
        i = int(rng.integers(0, t.n))
        if l[i] < 7:
            t.refine_key(k[i])
    c = t.centers()
    solid = (np.abs(c[:, 0] - 0.9) < 0.3) & (np.abs(c[:, 1] + 0.6) < 0.25) & (c[:, 2] < 1.2)
    for i in np.flatnonzero(solid):
        t.set_mask(int(i), 2)
    return t


def uniform_octree(orc, levels, max_level=6):
    t = orc.octree((0, 0, 0), (1, 1, 1), max_level)
    t.refine_uniform(levels)
    return t


def _morton_axis(k, axis):
    x = (k >> np.uint64(axis)) & np.uint64(0x1249249249249249)
    for sh, m in ((2, 0x10c30c30c30c30c3), (4, 0x100f00f00f00f00f), (8, 0x001f0000ff0000ff),
                  (16, 0x001f00000000ffff), (32, 0x1fffff)):
        x = (x ^ (x >> np.uint64(sh))) & np.uint64(m)
    return x


def mg_level_counts(keys, levels, max_level, extent, coarsest=2048, max_levels=16):
    """Cells per level of the octree multigrid hierarchy, restated with numpy:
    the target log2 box size S grows along the axes whose physical size is
    within 1.5x of the smallest; cells no larger than S group by their
    S-aligned box and a group that fills its box merges."""
    keys = np.asarray(keys, np.uint64)
    n = keys.size
    a = np.stack([_morton_axis(keys, q).astype(np.int64) for q in range(3)], axis=1)
    lg = np.repeat((max_level - np.asarray(levels, np.int64))[:, None], 3, axis=1)
    fine = np.asarray(extent, float) / float(1 << max_level)
    S = np.full(3, lg.min(), np.int64)
    counts = [n]
    while counts[-1] > coarsest and len(counts) < max_levels:
        merged = False
        while not (S >= max_level).all():
            h = fine * (2.0 ** S)
            S = np.where((h <= 1.5 * h.min()) & (S < max_level), S + 1, S)
            part = (lg <= S).all(axis=1)
            t = a & ~((np.int64(1) << S) - 1)
            tk = (t[:, 0] << 42) | (t[:, 1] << 21) | t[:, 2]
            tk = np.where(part, tk, -1)
            u, inv, cnt = np.unique(tk, return_inverse=True, return_counts=True)
            vol = np.bincount(inv, weights=np.where(part, 2.0 ** lg.sum(axis=1), 0.0))
            comp = (u >= 0) & (cnt > 1) & (vol == 2.0 ** S.sum())
            if not comp.any():
                continue
            cm = comp[inv]
            a = np.where(cm[:, None], t, a)
            lg = np.where(cm[:, None], S[None, :], lg)
            key = (a[:, 0] << 42) | (a[:, 1] << 21) | a[:, 2]
            uk, first = np.unique(key, return_index=True)
            a, lg = a[first], lg[first]
            counts.append(uk.size)
            merged = True
            break
        if not merged:
            break
    return counts


def coefficient_values(t):
    """Distinct face coefficients area / dist of the reference's octree
    operator (octree_poisson.cpp:24-32: the finer leaf's transverse face
    area, Octree::cellSize, over the distance of the Octree::cellCenter
    coordinates along the axis) over every fluid-fluid face of the tree."""
    _k, l, m = t.dump()
    _cnt, nbr = t.face_table()
    cen = t.centers()
    dxf = np.array(t.extent, float) / float(1 << t.max_level)
    sz = (1 << (t.max_level - l.astype(np.int64))).astype(float)
    vals = set()
    for d in range(6):
        a = d >> 1
        t1, t2 = (a + 1) % 3, (a + 2) % 3
        for q in range(4):
            j = nbr[:, d, q]
            i = np.flatnonzero((j >= 0) & (m == 0))
            j = j[i]
            keep = m[j] == 0
            i, j = i[keep], j[keep]
            fin = np.where(l[j] > l[i], j, i)
            area = (sz[fin] * dxf[t1]) * (sz[fin] * dxf[t2])
            vals.update((area / np.abs(cen[i, a] - cen[j, a])).tolist())
    return vals


def ladder_octree(orc, points=16, seed=3):
    """Non-dyadic box, base level 2, refined to level 14 around seeded points
    (each step refines the leaf holding the point): face coefficients between
    many level pairs at inexact centres, more than 256 distinct values
    (about 10k leaves)."""
    origin, extent, ml = (0.1, -0.37, 0.61), (0.93, 1.37, 0.71), 14
    t = orc.octree(origin, extent, ml)
    t.refine_uniform(2)
    rng = np.random.default_rng(seed)
    dxf = np.array(extent) / float(1 << ml)
    for _ in range(points):
        p = np.array([origin[a] + extent[a] * rng.uniform(0.05, 0.95) for a in range(3)])
        while True:
            k, l, _m = t.dump()
            half = (1 << (ml - l.astype(np.int64)))[:, None] * dxf[None, :] / 2
            i = int(np.flatnonzero(np.all(np.abs(t.centers() - p[None, :]) <= half, axis=1))[0])
            if l[i] >= ml:
                break
            t.refine_key(k[i])
    return t
