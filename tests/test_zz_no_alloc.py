"""nbl_poisson_solve_noalloc against nbl_poisson_solve through the C-ABI.

The no-alloc entry point stages b and x in the library's fixed per-GPU array
instead of two device allocations per call; everything else is the same solve,
so its results must be bit-identical to nbl_poisson_solve's.  Two calls in a
row share the array, the second on a smaller grid, and a grid above the array's
capacity is refused before anything is launched.
"""
import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu


def _block_channel(nx, ny, nz):
    """Channel with an outlet and a solid block (non-singular, solid cells)."""
    faces = [dict(type=t) for t in (nbl.BC_INLET, nbl.BC_OUTLET, nbl.BC_WALL, nbl.BC_WALL,
                                    nbl.BC_PERIODIC, nbl.BC_PERIODIC)]
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[nz // 4:nz // 2, : ny // 3, nx // 3:nx // 3 + 3] = 1
    return nbl.Grid(nx, ny, nz, 1.0 / ny, faces=faces, solid=solid.reshape(-1))


def _periodic(n):
    return cases.taylor_green(n)[0]


@pytest.mark.parametrize("precision", [nbl.PREC_FP64, nbl.PREC_MIXED])
def test_noalloc_solve_equals_solve(ctx, precision):
    print(f"noalloc solve, precision {precision}", flush=True)
    opts = nbl.solver_opts(precision=precision)
    # first a larger grid, then a smaller one: the second call reuses the array
    for g, seed in ((_block_channel(64, 32, 32), 11), (_periodic(16), 12), (_block_channel(24, 8, 8), 13)):
        b = cases.poisson_rhs(g, seed)
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-9, max_iters=200, opts=opts)
        try:
            x_ref, r_ref = ps.solve(b)
            x1, r1 = ps.solve_noalloc(b)
            x2, r2 = ps.solve_noalloc(b)
        finally:
            ps.close()
        assert r_ref.converged and r_ref.iterations > 0
        for x, r in ((x1, r1), (x2, r2)):
            assert (r.iterations, r.converged) == (r_ref.iterations, r_ref.converged)
            assert r.residual == r_ref.residual
            assert np.array_equal(x, x_ref)


def test_noalloc_jacobi_cg_equals_solve(ctx):
    g = _block_channel(32, 16, 16)
    b = cases.poisson_rhs(g, 21)
    ps = nbl.Poisson(ctx, g, kind=1, tol=1e-8, max_iters=2000)
    try:
        x_ref, r_ref = ps.solve(b)
        x, r = ps.solve_noalloc(b)
    finally:
        ps.close()
    assert (r.iterations, r.residual, r.converged) == (r_ref.iterations, r_ref.residual, r_ref.converged)
    assert np.array_equal(x, x_ref)


def test_noalloc_refuses_grids_above_capacity(ctx):
    g = nbl.Grid(128, 128, 129, 1.0 / 128)  # 2,113,536 cells > 2^21
    ps = nbl.Poisson(ctx, g, kind=1, tol=1e-6, max_iters=1)
    try:
        with pytest.raises(nbl.NblError) as e:
            ps.solve_noalloc(np.zeros(g.cells))
        assert e.value.code == -4 and "scratch" in str(e.value)
    finally:
        ps.close()
