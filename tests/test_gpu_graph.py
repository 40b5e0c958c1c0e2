"""Flexible-CG iterations replayed in pairs from a captured HIP graph against
the same iterations launched kernel by kernel, through the C-ABI: the graph
holds the same launches with the same arguments, so every solve returns the
same bits (solution, iteration count, residual).  Each solve replays the graph
several times with the solve's own eager launches (set-up, first iteration,
odd leftovers, final pass) and host synchronizations between replays; several
solves in a row reuse one capture.  Covers the per-sweep cycle (channel,
mixed precision, odd grid with solids), the two-kernels-per-level cycle
(singular Taylor-Green, BF16 form) and whole MAC steps."""
import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu


def grids():
    faces = [dict(type=nbl.BC_INLET, inlet_speed=1.0), dict(type=nbl.BC_OUTLET), dict(type=nbl.BC_WALL),
             dict(type=nbl.BC_WALL), dict(type=nbl.BC_WALL), dict(type=nbl.BC_OUTLET)]
    solid = np.zeros((20, 34, 50), np.uint8)
    solid[5:12, 0:9, 14:23] = 1
    return {"cube192": cases.cube_channel()[0], "tgv64": cases.taylor_green(64)[0],
            "odd50": nbl.Grid(50, 34, 20, 0.1, 0.12, 0.15, faces=faces, solid=solid)}


GRIDS = grids()
CASES = [("cube192", nbl.PREC_FP64), ("cube192", nbl.PREC_MIXED), ("tgv64", nbl.PREC_FP64),
         ("tgv64", nbl.PREC_MIXED_BF16), ("odd50", nbl.PREC_FP64)]


@pytest.mark.parametrize("name,prec", CASES, ids=[f"{n}-{p}" for n, p in CASES])
def test_graph_solves_bitwise(ctx, name, prec):
    g = GRIDS[name]
    out = {}
    for mode in (1, 2):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=200,
                         opts=nbl.solver_opts(precision=prec, iteration_graph=mode))
        try:
            out[mode] = [ps.solve(cases.poisson_rhs(g, s)) for s in (61, 62, 63)]
        finally:
            ps.close()
    for (xe, re), (xg, rg) in zip(out[1], out[2]):
        assert re.converged and rg.converged
        assert (re.iterations, re.residual) == (rg.iterations, rg.residual)
        assert np.array_equal(xe, xg)
    print(f"{name}: iterations {[r.iterations for _, r in out[2]]}", flush=True)


def test_graph_steps_bitwise(ctx):
    """Ten channel steps on 96x32x32 with graphs and without: equal states and reports."""
    g, cfg = cases.cube_channel(nx=96, ny=32, nz=32, res=8)
    st = cases.cube_channel_state(g)
    res = []
    for mode in (1, 2):
        run = nbl.Run(ctx, g, cfg, nbl.solver_opts(iteration_graph=mode))
        try:
            run.set_state(*st)
            reps = [run.step() for _ in range(10)]
            res.append((reps, run.get_state()))
        finally:
            run.close()
    (ra, sa), (rb, sb) = res
    for a, b in zip(ra, rb):
        assert a.as_dict() == b.as_dict()
    for fa, fb in zip(sa[:4], sb[:4]):
        assert np.array_equal(fa, fb)


def test_iteration_graph_validated(ctx):
    with pytest.raises(nbl.NblError):
        nbl.Poisson(ctx, GRIDS["odd50"], kind=0, opts=nbl.solver_opts(iteration_graph=3))
