"""Batched pressure solves (nbl_poisson_batch_*) through the C-ABI: problems on
one grid solved together, one launch per level kernel for all of them.

* Each problem's arithmetic is that of a batch of one: a 4-problem batch
  returns, per problem, the bits (solution, iterations, residual) the same
  problem gets alone in a 1-problem batch.
* Against the single-problem solver (whose reductions may fold inline, a
  different partial-sum tree) and the oracle's MgPoisson: iterations within
  one, solutions within rounding of the tolerance.
* Problems stop on their own: a zero right-hand side returns at once, a
  scaled one takes its unscaled twin's iterations, while the others go on.
Grids: the cube channel 96x32x32 (outlet: non-singular), the 32^3 periodic
Taylor-Green box and a 40x24x36 walled box (singular), and a 50x34x20 box
with a solid block, inlet and outlets (odd sizes); FP64 and FP32 levels.
"""
import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def grids():
    faces = [dict(type=nbl.BC_INLET, inlet_speed=1.0), dict(type=nbl.BC_OUTLET), dict(type=nbl.BC_WALL),
             dict(type=nbl.BC_WALL), dict(type=nbl.BC_WALL), dict(type=nbl.BC_OUTLET)]
    solid = np.zeros((20, 34, 50), np.uint8)
    solid[5:12, 0:9, 14:23] = 1
    return {"cube96": cases.cube_channel(nx=96, ny=32, nz=32, res=8)[0],
            "tgv32": cases.taylor_green(32)[0],
            "walls40": nbl.Grid(40, 24, 36, 1.0 / 40, faces=[dict(type=nbl.BC_WALL)] * 6),
            "odd50": nbl.Grid(50, 34, 20, 0.1, 0.12, 0.15, faces=faces, solid=solid)}


GRIDS = grids()
PRECS = [(nbl.PREC_FP64, "fp64"), (nbl.PREC_MIXED, "fp32")]


def batch_solve(ctx, g, bs, prec, tol=1e-10):
    with nbl.PoissonBatch(ctx, g, len(bs), tol=tol, max_iters=200, opts=nbl.solver_opts(precision=prec)) as pb:
        assert pb.size() == len(bs)
        return pb.solve(bs)


@pytest.mark.parametrize("prec,pname", PRECS, ids=[p for _, p in PRECS])
@pytest.mark.parametrize("name", sorted(GRIDS))
def test_batch_equals_batches_of_one(ctx, orc, name, prec, pname):
    g = GRIDS[name]
    bs = [cases.poisson_rhs(g, 70 + q) for q in range(4)]
    xs, reps = batch_solve(ctx, g, bs, prec)
    for q, b in enumerate(bs):
        (x1,), (r1,) = batch_solve(ctx, g, [b], prec)
        assert np.array_equal(xs[q], x1), q
        assert (reps[q].iterations, reps[q].residual, reps[q].converged) == (r1.iterations, r1.residual,
                                                                             r1.converged), q
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=200, opts=nbl.solver_opts(precision=prec))
    try:
        for q, b in enumerate(bs):
            xs1, rs1 = ps.solve(b)
            assert reps[q].converged and reps[q].residual <= 1e-10 * 1.01
            assert abs(reps[q].iterations - rs1.iterations) <= 1
            assert rel(xs[q], xs1) < 1e-9
    finally:
        ps.close()
    xo, ro = orc.mg(g, bs[0], 1e-10, 200)
    assert abs(reps[0].iterations - ro.iterations) <= (2 if prec == nbl.PREC_MIXED else 1)
    assert rel(xs[0], xo) < 1e-8
    print(f"{name} {pname}: iterations {[r.iterations for r in reps]}", flush=True)


@pytest.mark.parametrize("name", ["cube96", "walls40"])
def test_batch_problems_stop_on_their_own(ctx, name):
    g = GRIDS[name]
    b0, b1 = cases.poisson_rhs(g, 81), cases.poisson_rhs(g, 82)
    bs = [b0, np.zeros(g.cells), 1e6 * b1, b1]
    xs, reps = batch_solve(ctx, g, bs, nbl.PREC_FP64)
    assert reps[1].converged and reps[1].iterations == 0 and not np.any(xs[1])
    assert reps[2].iterations == reps[3].iterations
    assert rel(xs[2], 1e6 * xs[3]) < 1e-8
    (x0,), (r0,) = batch_solve(ctx, g, [b0], nbl.PREC_FP64)
    assert np.array_equal(xs[0], x0) and reps[0].iterations == r0.iterations
    assert all(r.converged for r in reps)


def test_batch_options_validated(ctx):
    g = GRIDS["cube96"]
    for opts in (nbl.solver_opts(loops=True), nbl.solver_opts(vcycle=nbl.VCYCLE_TILED),
                 nbl.solver_opts(precision=nbl.PREC_MIXED_BF16), nbl.solver_opts(launch_order=2),
                 nbl.solver_opts(coarse=nbl.COARSE_CG)):
        with pytest.raises(nbl.NblError):
            nbl.PoissonBatch(ctx, g, 2, opts=opts)
    with pytest.raises(nbl.NblError):
        nbl.PoissonBatch(ctx, g, 0)
    with pytest.raises(nbl.NblError):  # a single multigrid level
        nbl.PoissonBatch(ctx, nbl.Grid(3, 3, 3, 1.0 / 3), 2)
    with nbl.PoissonBatch(ctx, g, 2) as pb:
        single = nbl.Poisson.__new__(nbl.Poisson)
        single.lib, single.grid, single.h = pb.lib, g, pb.h
        with pytest.raises(nbl.NblError):
            single.solve(cases.poisson_rhs(g, 1))
    ps = nbl.Poisson(ctx, g, kind=0)
    try:
        batch = nbl.PoissonBatch.__new__(nbl.PoissonBatch)
        batch.lib, batch.grid, batch.h, batch.problems = ps.lib, g, ps.h, 1
        with pytest.raises(nbl.NblError):
            batch.solve([cases.poisson_rhs(g, 1)])
        assert batch.size() == 0
    finally:
        ps.close()
