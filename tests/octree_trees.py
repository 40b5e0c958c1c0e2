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
