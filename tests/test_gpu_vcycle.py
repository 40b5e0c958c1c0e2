"""The V(2,2) cycle in two kernels per level (LDS tiles with recomputed halos)
against the one-kernel-per-sweep cycle and the oracle, through the C-ABI.

* Systems without null-space projections: every cell value of the tiled cycle
  is the per-sweep kernels' expression in their order, so one preconditioner
  application is bit-identical; solves differ only through the partial-sum
  tree of the flexible CG's reductions (same iteration counts, agreement far
  inside the tolerance).
* Singular systems: the tiled cycle folds the per-sweep null-space projections
  into the ones that change values (exact in exact arithmetic), so results
  agree to rounding.
Grids have at least three multigrid levels (the tiled cycle's domain): 3-D and
2-D, odd sizes, solids, inlets and outlets, periodic boxes.
"""
import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def grids():
    out = {}
    out["tgv32"] = cases.taylor_green(32)[0]
    out["cube96"] = cases.cube_channel(nx=96, ny=32, nz=32, res=8)[0]
    g, _ = cases.lid_cavity(128, 100.0)
    out["lid128"] = g
    faces = [dict(type=nbl.BC_INLET, inlet_speed=1.0), dict(type=nbl.BC_OUTLET), dict(type=nbl.BC_WALL),
             dict(type=nbl.BC_WALL), dict(type=nbl.BC_WALL), dict(type=nbl.BC_OUTLET)]
    solid = np.zeros((20, 34, 50), np.uint8)
    solid[5:12, 0:9, 14:23] = 1
    out["odd50"] = nbl.Grid(50, 34, 20, 0.1, 0.12, 0.15, faces=faces, solid=solid)
    walls = nbl.Grid(40, 24, 36, 1.0 / 40, faces=[dict(type=nbl.BC_WALL)] * 6)
    out["walls40"] = walls
    return out


GRIDS = grids()
SINGULAR = {"tgv32", "lid128", "walls40"}


def solver(ctx, g, per_sweep, precision=nbl.PREC_FP64, tol=1e-10):
    return nbl.Poisson(ctx, g, kind=0, tol=tol, max_iters=200,
                       opts=nbl.solver_opts(precision=precision,
                                            vcycle=nbl.VCYCLE_PER_SWEEP if per_sweep else nbl.VCYCLE_TILED))


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_grids_have_three_levels(ctx, name):
    ps = solver(ctx, GRIDS[name], False)
    try:
        assert len(ps.levels()) >= 3
    finally:
        ps.close()


@pytest.mark.parametrize("precision", [nbl.PREC_FP64, nbl.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(set(GRIDS) - SINGULAR))
def test_tiled_preconditioner_bitwise(ctx, name, precision):
    g = GRIDS[name]
    r = cases.poisson_rhs(g, 41)
    zs = []
    for per_sweep in (True, False):
        ps = solver(ctx, g, per_sweep, precision)
        try:
            zs.append(ps.precondition(r))
        finally:
            ps.close()
    assert np.array_equal(zs[0], zs[1])


@pytest.mark.parametrize("precision", [nbl.PREC_FP64, nbl.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(GRIDS))
def test_tiled_solve_matches_per_sweep_and_oracle(ctx, orc, name, precision):
    g = GRIDS[name]
    b = cases.poisson_rhs(g, 43)
    out = []
    for per_sweep in (True, False):
        ps = solver(ctx, g, per_sweep, precision)
        try:
            out.append(ps.solve(b))
        finally:
            ps.close()
    (x1, r1), (x2, r2) = out
    print(f"{name}: iterations per-sweep {r1.iterations}, tiled {r2.iterations}; rel {rel(x2, x1):.2e}", flush=True)
    assert r1.converged and r2.converged and r2.residual <= 1e-10 * 1.01
    assert abs(r1.iterations - r2.iterations) <= (1 if name in SINGULAR else 0)
    assert rel(x2, x1) < 1e-8
    xo, ro = orc.mg(g, b, 1e-10, 200)
    assert abs(r2.iterations - ro.iterations) <= (2 if precision == nbl.PREC_MIXED else 1)
    assert rel(x2, xo) < 1e-8


@pytest.mark.parametrize("name", ["tgv32", "cube96"])
def test_tiled_solves_are_deterministic(ctx, name):
    g = GRIDS[name]
    b = cases.poisson_rhs(g, 47)
    ps = solver(ctx, g, False)
    try:
        xs = [ps.solve(b) for _ in range(4)]
    finally:
        ps.close()
    for x, r in xs[1:]:
        assert np.array_equal(x, xs[0][0]) and r.iterations == xs[0][1].iterations


def test_tiled_steps_track_per_sweep(ctx):
    """Cube channel 96x32x32 (three levels), WENO-5: five steps with either cycle."""
    g, cfg = cases.cube_channel(nx=96, ny=32, nz=32, res=8)
    st = cases.cube_channel_state(g)
    res = []
    for per_sweep in (True, False):
        run = nbl.Run(ctx, g, cfg, nbl.solver_opts(vcycle=nbl.VCYCLE_PER_SWEEP if per_sweep else nbl.VCYCLE_TILED))
        try:
            run.set_state(*st)
            reps = [run.step() for _ in range(5)]
            res.append((reps, run.get_state()))
        finally:
            run.close()
    (ra, sa), (rb, sb) = res
    for a, b in zip(ra, rb):
        assert a.poisson_iters == b.poisson_iters
        assert b.cd == pytest.approx(a.cd, rel=1e-10)
    for fa, fb in zip(sa[:4], sb[:4]):
        assert rel(fb, fa) < 1e-11


def test_tiled_taylor_green_steps_track_per_sweep(ctx):
    """64^3 periodic Taylor-Green vortex (singular): ten steps with either
    cycle agree to rounding."""
    g, cfg = cases.taylor_green(64)
    st = cases.taylor_green_state(g)
    res = []
    for vc in (nbl.VCYCLE_PER_SWEEP, nbl.VCYCLE_TILED):
        run = nbl.Run(ctx, g, cfg, nbl.solver_opts(vcycle=vc))
        try:
            run.set_state(*st)
            reps = [run.step() for _ in range(10)]
            res.append((reps, run.get_state()))
        finally:
            run.close()
    (ra, sa), (rb, sb) = res
    for a, b in zip(ra, rb):
        assert abs(a.poisson_iters - b.poisson_iters) <= 1
    for fa, fb in zip(sa[:4], sb[:4]):
        assert rel(fb, fa) < 1e-10


def test_vcycle_option_validated(ctx):
    g = GRIDS["tgv32"]
    with pytest.raises(nbl.NblError):
        nbl.Poisson(ctx, g, kind=0, opts=nbl.solver_opts(vcycle=3))


@pytest.mark.parametrize("name", ["tgv32", "cube96", "odd50"])
def test_bf16_scratch_reaches_fp64_tolerance(ctx, orc, name):
    """The BF16 form (the finest level's pre-smoothed iterate stored as BF16
    between the cycle's down and up kernels): the FP64 outer residual still
    reaches the tolerance, the solution agrees with the FP64 oracle to it, and
    the rounder preconditioner costs at most a few iterations."""
    g = GRIDS[name]
    b = cases.poisson_rhs(g, 53)
    out = {}
    for prec in (nbl.PREC_MIXED, nbl.PREC_MIXED_BF16):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=200, opts=nbl.solver_opts(prec))
        try:
            out[prec] = ps.solve(b)
        finally:
            ps.close()
    (x32, r32), (x16, r16) = out[nbl.PREC_MIXED], out[nbl.PREC_MIXED_BF16]
    xo, ro = orc.mg(g, b, 1e-10, 200)
    print(f"{name}: iterations fp64 {ro.iterations}, fp32 {r32.iterations}, bf16 {r16.iterations}", flush=True)
    assert r16.converged and r16.residual <= 1e-10 * 1.01
    assert r16.iterations <= r32.iterations + 4
    assert rel(x16, xo) < 1e-8


def test_bf16_needs_the_tiled_cycle(ctx):
    g, _ = cases.taylor_green(16)  # two levels: the two-kernels-per-level cycle does not apply
    with pytest.raises(nbl.NblError) as e:
        nbl.Poisson(ctx, g, kind=0, opts=nbl.solver_opts(nbl.PREC_MIXED_BF16))
    assert e.value.code == -1 and "two-kernels-per-level" in str(e.value)
    with pytest.raises(nbl.NblError):
        nbl.Poisson(ctx, GRIDS["cube96"], kind=0, opts=nbl.solver_opts(nbl.PREC_MIXED_BF16,
                                                                         vcycle=nbl.VCYCLE_PER_SWEEP))


@pytest.mark.parametrize("precision", [nbl.PREC_FP64, nbl.PREC_MIXED], ids=["fp64", "mixed"])
def test_tiled_cycle_beyond_2048_tiles(ctx, orc, precision):
    """A 100^3 periodic box has 2275 level-0 tiles (odd coarse levels 25, 13,
    7): the tiled cycle runs there, every reduction still folded inline, and
    solves like the per-sweep cycle and the oracle."""
    g, _ = cases.taylor_green(100)
    b = cases.poisson_rhs(g, 53)
    out = []
    for vc in (nbl.VCYCLE_PER_SWEEP, nbl.VCYCLE_TILED):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=200, opts=nbl.solver_opts(precision=precision, vcycle=vc))
        try:
            if vc == nbl.VCYCLE_TILED:
                ps.profile(4, 1)  # the tiled cycle's down kernel: raises unless that cycle is in use
            out.append(ps.solve(b))
        finally:
            ps.close()
    (x1, r1), (x2, r2) = out
    assert r1.converged and r2.converged and abs(r1.iterations - r2.iterations) <= 1
    assert rel(x2, x1) < 1e-8
    xo, ro = orc.mg(g, b, 1e-10, 200)
    assert abs(r2.iterations - ro.iterations) <= (2 if precision == nbl.PREC_MIXED else 1)
    assert rel(x2, xo) < 1e-8


def test_singular_per_sweep_beyond_inline_folds(ctx, orc):
    """A 128^3 periodic box has 4096 level-0 workgroups, beyond the inline
    folds: the flexible CG's reductions run as finalize kernels, z's fluid
    mean is finalized with beta and removed by the next A p where it forms p.
    FP64 and FP32 cycles solve like the oracle (iterations, solution)."""
    g, _ = cases.taylor_green(128)
    b = cases.poisson_rhs(g, 61)
    xo, ro = orc.mg(g, b, 1e-10, 200)
    for precision in (nbl.PREC_FP64, nbl.PREC_MIXED):
        ps = solver(ctx, g, True, precision)
        try:
            x, r = ps.solve(b)
        finally:
            ps.close()
        assert r.converged and abs(r.iterations - ro.iterations) <= (2 if precision == nbl.PREC_MIXED else 1)
        assert rel(x, xo) < 1e-8, precision
