"""The XCD launch orders of the level kernels (geo3n and blk3 in
nabla-ai_amd/csrc/mg.hip / common.hpp), restated on the CPU.  Slabs of planes
(slab = 1): the launch of 8 gx x gy x ceil(gz / 8) blocks maps onto the
gx x gy x gz tiles with every tile run exactly once, surplus blocks fall past
the last plane, linear block ids (the partial-sum index) are unique and below
the block count, and the blocks a round-robin dispatcher hands to one XCD
(linear id mod 8) cover one contiguous range of planes, visited in order.
Bands of rows (slab = 2): 8 gx x ceil(gy / 8) x gz blocks cover every tile
once, surplus blocks fall past the last row, and each XCD receives one
contiguous band of block rows of every plane, plane by plane."""
import pytest


def slab_blocks(gx, gy, gz):
    GX, GY, GZ = 8 * gx, gy, -(-gz // 8)
    for bz in range(GZ):
        for by in range(GY):
            for bx in range(GX):
                lin = bx + GX * (by + GY * bz)
                yield lin, (bx >> 3, by, (bx & 7) * GZ + bz)


@pytest.mark.parametrize("g", [(3, 8, 64), (2, 64, 256), (1, 4, 8), (1, 2, 9), (5, 3, 17), (1, 1, 1), (4, 4, 12)])
def test_slab_launch_covers_tiles_once(g):
    gx, gy, gz = g
    nblocks = 8 * gx * gy * -(-gz // 8)
    seen, lins = {}, set()
    per_xcd = {x: [] for x in range(8)}
    for lin, (tx, ty, tz) in slab_blocks(gx, gy, gz):
        assert lin not in lins and 0 <= lin < nblocks
        lins.add(lin)
        if tz < gz:
            seen[(tx, ty, tz)] = seen.get((tx, ty, tz), 0) + 1
            per_xcd[lin % 8].append((lin, tz))
    assert len(seen) == gx * gy * gz and set(seen.values()) == {1}
    for x, blocks in per_xcd.items():
        planes = [tz for _, tz in sorted(blocks)]
        assert planes == sorted(planes)  # dispatch order walks the slab plane by plane
        if planes:
            assert set(planes) == set(range(planes[0], planes[-1] + 1))  # one contiguous slab


def band_blocks(gx, gy, gz):
    GX, GY, GZ = 8 * gx, -(-gy // 8), gz
    for bz in range(GZ):
        for by in range(GY):
            for bx in range(GX):
                lin = bx + GX * (by + GY * bz)
                yield lin, (bx >> 3, (bx & 7) * GY + by, bz)


@pytest.mark.parametrize("g", [(2, 64, 256), (3, 8, 64), (1, 4, 8), (5, 13, 3), (1, 1, 1), (2, 80, 160), (4, 9, 2)])
def test_band_launch_covers_tiles_once(g):
    gx, gy, gz = g
    nblocks = 8 * gx * -(-gy // 8) * gz
    seen, lins = {}, set()
    per_xcd = {x: [] for x in range(8)}
    for lin, (tx, ty, tz) in band_blocks(gx, gy, gz):
        assert lin not in lins and 0 <= lin < nblocks
        lins.add(lin)
        assert tx < gx and tz < gz
        if ty < gy:
            seen[(tx, ty, tz)] = seen.get((tx, ty, tz), 0) + 1
            per_xcd[lin % 8].append((lin, ty, tz))
    assert len(seen) == gx * gy * gz and set(seen.values()) == {1}
    for x, blocks in per_xcd.items():
        if not blocks:
            continue
        rows = sorted({ty for _, ty, _ in blocks})
        assert rows == list(range(rows[0], rows[-1] + 1))  # one contiguous band of block rows
        planes = [tz for _, _, tz in sorted(blocks)]
        assert planes == sorted(planes)  # dispatch order walks the planes in order
