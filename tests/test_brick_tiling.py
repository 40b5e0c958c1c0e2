"""The brick decomposition of the WENO-5 right-hand side (nabla-ai_amd/csrc/ns.hip,
k_rhs_brick and brickArgs), restated on the CPU: every face of u, v and w is
written exactly once (by a brick or by the trailing zero-right-hand-side
blocks), every x/y flux product a face reads from LDS was written in the same
plane, and every stencil read stays inside the 3 ghost layers of the padded
copies with at most one index out of range (edge and corner ghosts are never
resolved, so a read there would differ from the reference's accessor).  It
checks the tiling design against these rules; the GPU tests check the kernel's
results bit for bit."""
import numpy as np
import pytest

H = 3            # ghost layers of the padded copies
BX, BY = 64, 4   # faces along i, rows along j per brick


def brick_args(nx, ny, nz, planes):
    def tile(p):
        comps, tot = [], 0
        for c in range(3):
            ox, oy, oz = int(c == 0), int(c == 1), int(c == 2)
            ex, ey, ez = nx - (c == 0), ny - (c == 1), nz - (c == 2)
            tx, ty, tz = max(1, -(-ex // BX)), max(1, -(-ey // BY)), max(1, -(-ez // p))
            blocks = tx * ty * tz if min(ex, ey, ez) > 0 else 0
            comps.append((ox, oy, oz, ex, ey, ez, tx, ty, tz, blocks))
            tot += blocks
        return comps, tot
    if planes > 0:
        return tile(planes)[0], planes
    for p in (4, 2, 1):
        comps, tot = tile(p)
        if tot >= 4096:
            break
    return comps, p


def check_grid(nx, ny, nz, planes):
    dims = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    flat = nz == 1
    comps, P = brick_args(nx, ny, nz, planes)

    def read(comp, ii, jj, kk):
        X, Y, Z = dims[comp]
        for a, n in ((ii, X), (jj, Y), (kk, Z)):
            assert a.size == 0 or (a.min() >= -H and a.max() < n + H), "read outside the ghost layers"
        out = ((ii < 0) | (ii >= X)).astype(int) + ((jj < 0) | (jj >= Y)) + ((kk < 0) | (kk >= Z))
        assert not np.any(out > 1), "edge or corner ghost read"

    def flux(c, d, ii, jj, kk):
        e = [0, 0, 0]
        e[d] = 1
        for m in range(-3, 3):
            read(c, ii + m * e[0], jj + m * e[1], kk + m * e[2])
        ec = [0, 0, 0]
        ec[c] = 1
        read(d, ii, jj, kk)
        read(d, ii - ec[0], jj - ec[1], kk - ec[2])

    written = [np.zeros(d[::-1], np.int32) for d in dims]
    for c in range(3):
        ox, oy, oz, ex, ey, ez, tx, ty, tz, blocks = comps[c]
        for bt in range(blocks):
            bi, bj, bk = bt % tx, (bt // tx) % ty, bt // (tx * ty)
            i0, j0, k0 = ox + bi * BX, oy + bj * BY, oz + bk * P
            xe, ye, ke = ox + ex, oy + ey, min(k0 + P, oz + ez)
            l, r = np.meshgrid(np.arange(BX), np.arange(BY))
            i, j = i0 + l, j0 + r
            inx, iny = i < xe, j < ye
            inn = inx & iny
            if not flat:
                flux(c, 2, i[inn], j[inn], np.full(inn.sum(), k0))
            for k in range(k0, ke):
                lx = np.zeros((BY, BX + 1), bool)
                ly = np.zeros((BY + 1, BX), bool)
                m = (i <= xe) & iny
                flux(c, 0, i[m], j[m], np.full(m.sum(), k))
                lx[:, :BX][m] = True
                m = inx & (j <= ye)
                flux(c, 1, i[m], j[m], np.full(m.sum(), k))
                ly[:BY][m] = True
                if j0 + BY <= ye:
                    m = inx[0]
                    flux(c, 1, i[0][m], np.full(m.sum(), j0 + BY), np.full(m.sum(), k))
                    ly[BY][m] = True
                if i0 + BX <= xe:
                    rows = np.arange(BY)
                    m = j0 + rows < ye
                    flux(c, 0, np.full(m.sum(), i0 + BX), (j0 + rows)[m], np.full(m.sum(), k))
                    lx[rows[m], BX] = True
                if not flat:
                    flux(c, 2, i[inn], j[inn], np.full(inn.sum(), k + 1))
                rr, ll = np.nonzero(inn)
                assert lx[rr, ll].all() and lx[rr, ll + 1].all(), "x flux read before it was written"
                assert ly[rr, ll].all() and ly[rr + 1, ll].all(), "y flux read before it was written"
                ii, jj = i[inn], j[inn]
                kk = np.full(len(ii), k)
                for e in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    if e[2] and flat:
                        continue
                    read(c, ii + e[0], jj + e[1], kk + e[2])
                    read(c, ii - e[0], jj - e[1], kk - e[2])
                np.add.at(written[c], (kk, jj, ii), 1)
    edge = [2 * ny * nz, 2 * nx * nz, 2 * nx * ny]
    for c in range(3):
        q = np.arange(edge[c])
        s, h = q & 1, q >> 1
        if c == 0:
            i, j, k = np.where(s, nx, 0), h % ny, h // ny
        elif c == 1:
            i, j, k = h % nx, np.where(s, ny, 0), h // nx
        else:
            i, j, k = h % nx, h // nx, np.where(s, nz, 0)
        read(c, i, j, k)
        np.add.at(written[c], (k, j, i), 1)
    for c in range(3):
        assert np.all(written[c] == 1), f"component {c}: faces written {np.unique(written[c])} times"


@pytest.mark.parametrize("shape", [(12, 12, 1), (16, 8, 1), (12, 8, 6), (10, 9, 7), (130, 9, 11), (65, 5, 3),
                                   (1, 1, 1), (2, 3, 1), (64, 4, 17), (70, 5, 1)])
@pytest.mark.parametrize("planes", [0, 1, 3, 16])
def test_brick_tiling(shape, planes):
    check_grid(*shape, planes)
