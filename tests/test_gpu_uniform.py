"""GPU parity of the uniform MAC path against the oracle (oracle/cpu), through
the C-ABI.

Bars:
* per-element kernels (applyA, diagonal, one smoother sweep, restriction,
  prolongation, the momentum right-hand side) are bit-identical: the kernels
  evaluate the reference's expressions in the reference's order with FP
  contraction off;
* solves and time steps contain reductions whose summation order differs from
  the reference's serial loops, so they are compared with tolerances: the
  same iteration counts, solutions within rel-L2 1e-9 (solve tolerance
  1e-10), fields within rel-L2 1e-10 per step sequence.
"""
import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    d = np.linalg.norm(a - b)
    s = np.linalg.norm(b)
    return d / s if s > 0 else d


def grids():
    out = []
    g, _ = cases.lid_cavity(12, 100.0)
    out.append(("lid2d", g))
    g, _ = cases.channel(16, 8, 100.0)
    out.append(("chan2d", g))
    g, _ = cases.taylor_green(8)
    out.append(("tgv8", g))
    faces = [dict(type=nbl.BC_INLET, inlet_speed=1.0), dict(type=nbl.BC_OUTLET),
             dict(type=nbl.BC_WALL), dict(type=nbl.BC_WALL, wall_velocity=(0.3, 0.0, 0.1)),
             dict(type=nbl.BC_PERIODIC), dict(type=nbl.BC_PERIODIC)]
    solid = np.zeros((6, 8, 12), np.uint8)
    solid[2:4, 0:3, 4:7] = 1
    out.append(("box3d", nbl.Grid(12, 8, 6, 0.1, 0.12, 0.15, faces=faces, solid=solid)))
    faces2 = [dict(type=nbl.BC_WALL), dict(type=nbl.BC_WALL),
              dict(type=nbl.BC_INLET, profile=nbl.INLET_PARABOLIC, inlet_speed=0.7, inlet_axis=1),
              dict(type=nbl.BC_OUTLET), dict(type=nbl.BC_WALL), dict(type=nbl.BC_OUTLET)]
    out.append(("odd3d", nbl.Grid(10, 9, 7, 0.2, faces=faces2)))
    g, _ = cases.cube_channel(nx=48, ny=16, nz=16, res=4)
    out.append(("cube48", g))
    return out


GRIDS = grids()
IDS = [n for n, _ in GRIDS]


@pytest.mark.parametrize("name,g", GRIDS, ids=IDS)
def test_operator_pieces_bitwise(ctx, orc, name, g):
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=100)
    try:
        x = cases.normal(5, g.cells)
        b = cases.normal(6, g.cells)
        assert np.array_equal(ps.apply(x), orc.apply(g, x))
        assert np.array_equal(ps.diagonal(), orc.diag(g))
        assert ps.levels() == orc.levels(g)
        assert np.array_equal(ps.smooth_once(b, x), orc.smooth(g, b, x))
        if len(ps.levels()) > 1:
            assert np.array_equal(ps.restrict(b), orc.restrict(g, b))
            ec = cases.normal(8, int(np.prod(ps.levels()[1])))
            assert np.array_equal(ps.prolong(ec, x), orc.prolong(g, ec, x))
    finally:
        ps.close()


@pytest.mark.parametrize("coarse", [nbl.COARSE_DIRECT, nbl.COARSE_CG], ids=["direct", "cg"])
@pytest.mark.parametrize("name,g", GRIDS, ids=IDS)
def test_mg_pcg_matches_oracle(ctx, orc, name, g, coarse):
    b = cases.poisson_rhs(g, 11)
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=100, opts=nbl.solver_opts(coarse=coarse))
    try:
        x, rep = ps.solve(b)
        xo, ro = orc.mg(g, b, 1e-10, 100)
        assert rep.converged and ro.converged
        assert abs(rep.iterations - ro.iterations) <= 1
        assert rep.residual <= 1e-10 * 1.01
        assert rel(x, xo) < 1e-9
        # a second solve on the same object reuses the hierarchy and agrees
        x2, rep2 = ps.solve(b)
        assert np.array_equal(x2, x) and rep2.iterations == rep.iterations
    finally:
        ps.close()


@pytest.mark.parametrize("name,g", GRIDS, ids=IDS)
def test_reference_loops_match_oracle(ctx, orc, name, g):
    """The unfused baseline (one kernel per loop of multigrid.cpp with its
    CgPoisson(1e-6, 500) coarsest solve) converges as the reference does: the
    iteration count within 1 of the oracle's, solution within rel-L2 1e-9, and
    it agrees with the fused path to the same bar."""
    b = cases.poisson_rhs(g, 11)
    xo, ro = orc.mg(g, b, 1e-10, 100)
    out = []
    for loops in (True, False):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=100, opts=nbl.solver_opts(loops=loops))
        try:
            out.append(ps.solve(b))
        finally:
            ps.close()
    (xl, rl), (xf, rf) = out
    assert rl.converged and ro.converged
    assert abs(rl.iterations - ro.iterations) <= 1
    assert rl.residual <= 1e-10 * 1.01
    assert rel(xl, xo) < 1e-9 and rel(xl, xf) < 1e-9


def test_reference_loops_step_tracks_fused(ctx):
    """Ten cube-channel steps with the unfused baseline solve agree with the
    fused FP64 path (fields within rel-L2 1e-10)."""
    g, cfg = cases.cube_channel(nx=48, ny=16, nz=16, res=4)
    st = cases.cube_channel_state(g)
    res = []
    for loops in (True, False):
        run = nbl.Run(ctx, g, cfg, nbl.solver_opts(loops=loops))
        try:
            run.set_state(*st)
            for _ in range(10):
                rep = run.step()
                assert rep.poisson_residual <= cfg.poisson_tol
            res.append(run.get_state())
        finally:
            run.close()
    for a, b in zip(res[0][:4], res[1][:4]):
        assert rel(a, b) < 1e-10


def test_reference_loops_rejects_mixed(ctx):
    g, _ = cases.taylor_green(8)
    with pytest.raises(Exception):
        nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=10, opts=nbl.solver_opts(precision=nbl.PREC_MIXED, loops=True))


@pytest.mark.parametrize("prec", [nbl.PREC_FP64, nbl.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", ["cube192", "tgv64", "odd3d"])
def test_cells_per_thread_do_not_change_results(ctx, name, prec):
    """The level kernels process 1, 2 or 4 consecutive cells per thread with
    vector loads; per-cell arithmetic is the same, only the partial-sum tree
    of the reductions changes: identical iteration counts, solutions within
    rounding (rel-L2 1e-12 FP64, 1e-6 with the FP32 V-cycle)."""
    if name == "odd3d":
        g = dict(GRIDS)["odd3d"]
    else:
        g = cases.cube_channel()[0] if name == "cube192" else cases.taylor_green(64)[0]
    b = cases.poisson_rhs(g, 17)
    out = []
    for cpt in (1, 2, 4):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=100,
                         opts=nbl.solver_opts(precision=prec, cells_per_thread=cpt))
        try:
            out.append(ps.solve(b))
        finally:
            ps.close()
    (x1, r1), (x2, r2), (x4, r4) = out
    assert r1.converged and r2.converged and r4.converged
    assert r1.iterations == r2.iterations == r4.iterations
    tol = 1e-12 if prec == nbl.PREC_FP64 else 1e-6
    assert rel(x2, x1) < tol and rel(x4, x1) < tol


def _order_grid(name):
    if name == "odd3d":
        return dict(GRIDS)["odd3d"]
    if name == "cube192":
        return cases.cube_channel()[0]
    if name == "tgv64":
        return cases.taylor_green(64)[0]
    # 256 x 256 x 128 (8.4M cells, bands of rows by default) with an outlet:
    # non-singular, so the preconditioner has no reductions
    faces = [dict(type=nbl.BC_WALL)] * 5 + [dict(type=nbl.BC_OUTLET)]
    return nbl.Grid(256, 256, 128, 1.0 / 256, faces=faces)


@pytest.mark.parametrize("name", ["cube192", "odd3d", "tgv64", "big256x256x128"])
def test_launch_orders_do_not_change_results(ctx, name):
    """Plain order, slabs of planes and bands of rows per XCD run the same
    per-cell arithmetic: the preconditioner of a non-singular system (no
    reductions) is bit-identical across orders; solves have identical
    iteration counts and agree to rounding (only the partial-sum tree of the
    reductions changes)."""
    g = _order_grid(name)
    b = cases.poisson_rhs(g, 23)
    zs, out = [], []
    for order in (1, 2, 3):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=100, opts=nbl.solver_opts(launch_order=order))
        try:
            if name in ("cube192", "big256x256x128"):
                zs.append(ps.precondition(b))
            out.append(ps.solve(b))
        finally:
            ps.close()
    for z in zs[1:]:
        assert np.array_equal(z, zs[0])
    (x1, r1), (x2, r2), (x3, r3) = out
    print(f"{name}: iterations {r1.iterations} {r2.iterations} {r3.iterations}", flush=True)
    assert r1.converged and r2.converged and r3.converged
    assert r1.iterations == r2.iterations == r3.iterations
    assert rel(x2, x1) < 1e-12 and rel(x3, x1) < 1e-12


def test_launch_order_validated(ctx):
    g = cases.taylor_green(16)[0]
    for bad in (-1, 4):
        with pytest.raises(nbl.NblError):
            nbl.Poisson(ctx, g, kind=0, opts=nbl.solver_opts(launch_order=bad))


@pytest.mark.parametrize("name", ["cube192", "tgv64"])
def test_direct_coarse_matches_cg_coarse(ctx, name):
    """The exact coarsest-level solve and the reference's CG(1e-6, 500) there
    give the same flexible-CG iteration counts and solutions that agree far
    inside the outer tolerance (1e-8 solve, agreement <= 1e-11)."""
    g = cases.cube_channel()[0] if name == "cube192" else cases.taylor_green(64)[0]
    out = []
    for coarse in (nbl.COARSE_DIRECT, nbl.COARSE_CG):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=100, opts=nbl.solver_opts(coarse=coarse))
        try:
            out.append([ps.solve(cases.poisson_rhs(g, s)) for s in (1, 2, 3)])
        finally:
            ps.close()
    for (xd, rd), (xc, rc) in zip(*out):
        assert rd.converged and rc.converged and rd.iterations == rc.iterations
        assert rel(xd, xc) < 1e-11


@pytest.mark.parametrize("name,g", GRIDS[:4], ids=IDS[:4])
def test_cg_jacobi_matches_oracle(ctx, orc, name, g):
    b = cases.poisson_rhs(g, 13)
    ps = nbl.Poisson(ctx, g, kind=1, tol=1e-10, max_iters=3000)
    try:
        x, rep = ps.solve(b)
        xo, ro = orc.cg(g, b, 1e-10, 3000)
        assert abs(rep.iterations - ro.iterations) <= 2
        assert rel(x, xo) < 1e-8
        assert rep.residual < 1e-9
    finally:
        ps.close()


def test_mg_manufactured_solid_block(ctx):
    """core/tests/test_flow.cpp:63-117: 32x16x16 inlet/outlet/walls/periodic with
    a wall-mounted block; x_exact random on fluid; <= 50 cycles, err < 1e-6."""
    faces = [dict(type=nbl.BC_INLET), dict(type=nbl.BC_OUTLET), dict(type=nbl.BC_WALL),
             dict(type=nbl.BC_WALL), dict(type=nbl.BC_PERIODIC), dict(type=nbl.BC_PERIODIC)]
    solid = np.zeros((16, 16, 32), np.uint8)
    solid[6:10, 0:6, 12:18] = 1
    g = nbl.Grid(32, 16, 16, 1.0 / 16, faces=faces, solid=solid)
    xe = 2.0 * cases.uniform01(12345, g.cells) - 1.0
    xe[g.solid == 1] = 0.0
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=100)
    b = ps.apply(xe)
    x, rep = ps.solve(b)
    ps.close()
    assert rep.converged and rep.residual < 1e-8 and rep.iterations <= 50
    assert np.sqrt(np.mean((x - xe) ** 2)) < 1e-6


@pytest.mark.parametrize("conv", [nbl.CONV_CENTRAL, nbl.CONV_WENO5], ids=["central", "weno5"])
@pytest.mark.parametrize("name,g", GRIDS, ids=IDS)
def test_rhs_bitwise(ctx, orc, name, g, conv):
    cfg = nbl.flow_cfg(reynolds=200.0, convection=conv, body_force=(0.01, -0.02, 0.003))
    st = [cases.normal(s, n) * 0.3 for s, n in zip((1, 2, 3, 4), g.shapes)]
    run = nbl.Run(ctx, g, cfg)
    o = orc.ns(g, cfg)
    try:
        run.set_state(*st)
        o.set_state(*st)
        for a, b in zip(run.rhs(), o.rhs()):
            assert np.array_equal(a, b)
        for comp in range(3):
            assert np.array_equal(run.cell_velocity(comp), o.cell_velocity(comp))
    finally:
        run.close()


def brick_grids():
    faces = [dict(type=nbl.BC_INLET, inlet_speed=1.0), dict(type=nbl.BC_OUTLET),
             dict(type=nbl.BC_WALL, wall_velocity=(0.2, 0.0, -0.1)), dict(type=nbl.BC_WALL),
             dict(type=nbl.BC_PERIODIC), dict(type=nbl.BC_PERIODIC)]
    solid = np.zeros((11, 9, 130), np.uint8)
    solid[3:8, 0:4, 60:75] = 1
    solid[:, 8, 127:] = 1
    out = [("wide3d", nbl.Grid(130, 9, 11, 0.05, 0.07, 0.09, faces=faces, solid=solid))]
    g, _ = cases.channel(70, 5, 100.0)
    out.append(("wide2d", g))
    out.append(("cube192", cases.cube_channel()[0]))
    return out


@pytest.mark.parametrize("name,g", brick_grids(), ids=[n for n, _ in brick_grids()])
def test_rhs_bricks_bitwise(ctx, orc, name, g):
    """The WENO-5 right-hand side runs in 64 x 4 face bricks marching along k
    and shares each face's flux products with its neighbours.  On grids with
    several bricks per axis, ragged last bricks, solid cells and 2-D, for the
    automatic march and for marches of 1, 3 and 16 planes (the last brick of a
    column shorter than the rest), it is bit-identical to the oracle's
    per-face evaluation (ns_solver.cpp:311-473)."""
    cfg = nbl.flow_cfg(reynolds=300.0, convection=nbl.CONV_WENO5, body_force=(0.02, 0.01, -0.005))
    st = [cases.normal(s, n) * 0.4 for s, n in zip((21, 22, 23, 24), g.shapes)]
    o = orc.ns(g, cfg)
    o.set_state(*st)
    want = o.rhs()
    for planes in (0, 1, 3, 16):
        run = nbl.Run(ctx, g, cfg, nbl.solver_opts(rhs_planes=planes))
        try:
            run.set_state(*st)
            for a, b in zip(run.rhs(), want):
                assert np.array_equal(a, b), planes
        finally:
            run.close()


def test_rhs_planes_rejected(ctx):
    g, cfg = cases.lid_cavity(8, 100.0)
    with pytest.raises(nbl.NblError):
        nbl.Run(ctx, g, cfg, nbl.solver_opts(rhs_planes=17))


@pytest.mark.parametrize("conv", [nbl.CONV_CENTRAL, nbl.CONV_WENO5], ids=["central", "weno5"])
@pytest.mark.parametrize("name,g", GRIDS, ids=IDS)
def test_steps_match_oracle(ctx, orc, name, g, conv):
    cfg = nbl.flow_cfg(reynolds=200.0, convection=conv, poisson_tol=1e-10, cfl=0.5,
                       body_force=(0.01, 0.0, 0.0))
    st = [cases.normal(s, n) * 0.3 for s, n in zip((1, 2, 3, 4), g.shapes)]
    run = nbl.Run(ctx, g, cfg)
    o = orc.ns(g, cfg)
    try:
        run.set_state(*st)
        o.set_state(*st)
        assert run.compute_dt() == o.compute_dt()
        for _ in range(3):
            r, q = run.step(), o.step()
            assert r.step == q.step and r.poisson_iters == q.poisson_iters
            assert r.dt == pytest.approx(q.dt, rel=1e-12)
            assert r.t == pytest.approx(q.t, rel=1e-12)
            assert r.continuity_residual == pytest.approx(q.continuity_residual, rel=1e-3, abs=1e-13)
            assert r.momentum_residual == pytest.approx(q.momentum_residual, rel=1e-9)
            assert r.mass_error == pytest.approx(q.mass_error, rel=1e-6, abs=1e-13)
            assert r.cd == pytest.approx(q.cd, rel=1e-9, abs=1e-12)
            assert r.cl == pytest.approx(q.cl, rel=1e-9, abs=1e-12)
            assert r.accepted == q.accepted
        a, b = run.get_state(), o.get_state()
        for fa, fb in zip(a[:4], b[:4]):
            assert rel(fa, fb) < 1e-10
        assert a[4] == b[4]
    finally:
        run.close()


def test_initialize_matches_oracle(ctx, orc):
    g, cfg = cases.wall_mounted_cube(4, 500.0)
    run = nbl.Run(ctx, g, cfg)
    o = orc.ns(g, cfg)
    run.initialize()
    o.initialize()
    for fa, fb in zip(run.get_state()[:4], o.get_state()[:4]):
        assert np.array_equal(fa, fb)
    run.close()


def test_set_state_rejects_wrong_sizes(ctx):
    g, cfg = cases.lid_cavity(8, 100.0)
    run = nbl.Run(ctx, g, cfg)
    su, sv, sw, sp = g.shapes
    with pytest.raises(nbl.NblError):
        run.set_state(np.zeros(su + 1), np.zeros(sv), np.zeros(sw), np.zeros(sp))
    run.close()


def test_tgv16_ten_steps(ctx, orc):
    g, cfg = cases.taylor_green(16)
    st = cases.taylor_green_state(g)
    run = nbl.Run(ctx, g, cfg)
    o = orc.ns(g, cfg)
    run.set_state(*st)
    o.set_state(*st)
    for _ in range(10):
        r, q = run.step(), o.step()
        assert r.poisson_iters == q.poisson_iters
    a, b = run.get_state(), o.get_state()
    for fa, fb in zip(a[:4], b[:4]):
        assert rel(fa, fb) < 1e-10
    run.close()


@pytest.mark.parametrize("name,g", GRIDS, ids=IDS)
def test_mixed_precision_reaches_fp64_tolerance(ctx, orc, name, g):
    """FP32 V-cycle inside the FP64 flexible CG: the FP64 residual still reaches
    the requested tolerance, and the solution agrees with the FP64 oracle to it."""
    b = cases.poisson_rhs(g, 17)
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=200, opts=nbl.solver_opts(nbl.PREC_MIXED))
    try:
        x, rep = ps.solve(b)
    finally:
        ps.close()
    xo, ro = orc.mg(g, b, 1e-10, 200)
    assert rep.converged and rep.residual <= 1e-10 * 1.01
    assert rep.iterations <= ro.iterations + 3
    assert rel(x, xo) < 1e-8


def test_mixed_steps_track_fp64(ctx):
    g, cfg = cases.cube_channel(nx=48, ny=16, nz=16, res=4)
    cfg.poisson_tol = 1e-8
    st = cases.cube_channel_state(g)
    out = []
    for prec in (nbl.PREC_FP64, nbl.PREC_MIXED):
        run = nbl.Run(ctx, g, cfg, nbl.solver_opts(prec))
        run.set_state(*st)
        reps = [run.step() for _ in range(5)]
        out.append((reps, run.get_state()))
        run.close()
    (r64, s64), (r32, s32) = out
    for a, b in zip(r64, r32):
        assert b.poisson_residual <= 1e-8 and a.poisson_residual <= 1e-8
        assert b.cd == pytest.approx(a.cd, rel=1e-6)
    for fa, fb in zip(s64[:3], s32[:3]):
        assert rel(fb, fa) < 1e-7


@pytest.mark.parametrize("n", [24, 32])
def test_singular_mg_is_deterministic(ctx, n):
    """Repeated solves of a singular system (3-D restriction blocks, lazy mean
    projections) return the same bits; a partial written by two threads of a
    3-D block once made these differ in the last place."""
    g, _ = cases.taylor_green(n)
    b = cases.poisson_rhs(g, 3)
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=100)
    try:
        xs = [ps.solve(b)[0] for _ in range(8)]
        zs = [ps.precondition(b - b.mean()) for _ in range(4)]
    finally:
        ps.close()
    for x in xs[1:]:
        assert np.array_equal(x, xs[0])
    for z in zs[1:]:
        assert np.array_equal(z, zs[0])


@pytest.mark.parametrize("cpt", [2, 4])
def test_mixed_singular_vector_paths_match_oracle(ctx, orc, cpt):
    """The mixed form on a singular grid wide enough for 2 and 4 cells per
    thread: the kernel that removes r_new's mean also writes the FP32 copy the
    V-cycle starts from with 8- and 16-byte stores.  Against the FP64 oracle."""
    g, _ = cases.taylor_green(32)
    b = cases.poisson_rhs(g, 29)
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=200,
                     opts=nbl.solver_opts(nbl.PREC_MIXED, cells_per_thread=cpt))
    try:
        x, rep = ps.solve(b)
    finally:
        ps.close()
    xo, ro = orc.mg(g, b, 1e-10, 200)
    assert rep.converged and rep.residual <= 1e-10 * 1.01
    assert abs(rep.iterations - ro.iterations) <= 2
    assert rel(x, xo) < 1e-8

