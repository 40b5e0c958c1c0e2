"""The flow step on octree leaves (nbl_octree_flow_*) through the C-ABI.

The reference has no octree momentum step (its design record plans one,
docs/adr/0008-octree-native-flow-solve.md:68-86), so the step is pinned by what
the discretization must do, on trees with 2:1 level jumps:
* free-stream preservation: uniform inflow between free-slip walls stays
  uniform to round-off across level jumps (WENO-5 and upwind faces);
* the projection: after a step every fluid leaf's net face flux vanishes to
  the pressure solve's tolerance, inflow equals outflow, and in a closed box
  kinetic energy decays (upwind convection, viscosity, no forcing);
* plane Poiseuille flow: with the parabolic inflow the analytic profile is a
  steady state; after 100 steps the profile stays within 2% (rel. L2) of it
  and the core pressure gradient within 2% of the analytic -12 nu U / H^2;
* solid leaves carry no velocity; repeated runs are bit-identical.
"""
import numpy as np
import pytest

import nbl

pytestmark = pytest.mark.gpu


def channel_tree(orc, max_level=6, base=4, walls_refined=True):
    t = orc.octree((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), max_level)
    t.refine_uniform(base)
    if walls_refined:  # one level finer along both y walls: 2:1 jumps at y = 1/8 and 7/8
        t.refine_box((0.0, 0.0, 0.0), (1.0, 0.12, 1.0), 1)
        t.refine_box((0.0, 0.88, 0.0), (1.0, 1.0, 1.0), 1)
    return t


def jump_tree(orc):
    t = orc.octree((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), 6)
    t.refine_uniform(3)
    t.refine_box((0.6, 0.3, 0.3), (1.2, 0.7, 0.7), 2)  # two nested refinements in the middle
    return t


def make_flow(ctx, t, **kw):
    k, l, m = t.dump()
    d = nbl.Octree(ctx, k, l, m, t.origin, t.extent, t.max_level)
    return d, nbl.OctreeFlow(d, nbl.oct_flow_cfg(**kw))


SLIP_CHANNEL = (nbl.OCT_INLET, nbl.OCT_OUTLET, nbl.OCT_SLIP, nbl.OCT_SLIP, nbl.OCT_SLIP, nbl.OCT_SLIP)


@pytest.mark.parametrize("weno", [0, 1], ids=["upwind", "weno5"])
def test_free_stream_is_preserved(ctx, orc, weno):
    t = jump_tree(orc)
    d, f = make_flow(ctx, t, nu=0.01, cfl=0.5, inlet_speed=1.0, face=SLIP_CHANNEL, convection=weno,
                     poisson_tol=1e-10)
    try:
        n = t.n
        f.set_state(np.ones(n), np.zeros(n), np.zeros(n))
        for _ in range(5):
            r = f.step()
            assert abs(r.inflow - r.outflow) <= 1e-12 * r.inflow
        u, v, w, p = f.get_state()
        assert np.abs(u - 1.0).max() < 1e-12 and np.abs(v).max() < 1e-12 and np.abs(w).max() < 1e-12
        assert np.abs(f.divergence()).max() < 1e-10
    finally:
        f.close()
        d.close()


def smooth_field(c):
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    return (np.sin(np.pi * x) * np.cos(np.pi * y) * np.sin(2 * np.pi * z),
            -np.cos(np.pi * x) * np.sin(2 * np.pi * y) * np.sin(np.pi * z),
            0.3 * np.sin(3 * np.pi * x) * np.sin(np.pi * y))


def test_projection_in_a_closed_box(ctx, orc):
    t = jump_tree(orc)
    walls = (nbl.OCT_WALL,) * 6
    d, f = make_flow(ctx, t, nu=0.02, cfl=0.4, face=walls, convection=0, poisson_tol=1e-11)
    try:
        u0, v0, w0 = smooth_field(t.centers())
        f.set_state(u0, v0, w0)
        ke = []
        for _ in range(8):
            r = f.step()
            assert r.inflow == 0.0 and abs(r.outflow) == 0.0
            assert r.divergence_max < 1e-8, r.divergence_max
            ke.append(r.kinetic_energy)
        assert all(b < a for a, b in zip(ke, ke[1:])), ke
        div = f.divergence()
        assert np.abs(div).max() < 1e-8
        # the net flux of every leaf vanishes, so the flux through any union of leaves does too
        vol = np.prod(np.array(t.extent)) / 8.0 ** np.asarray(t.dump()[1], float)
        assert abs(np.sum(div * vol)) < 1e-12
    finally:
        f.close()
        d.close()


def test_poiseuille_profile(ctx, orc):
    t = channel_tree(orc)
    nu, U, H = 0.1, 1.0, 1.0
    faces = (nbl.OCT_INLET, nbl.OCT_OUTLET, nbl.OCT_WALL, nbl.OCT_WALL, nbl.OCT_SLIP, nbl.OCT_SLIP)
    d, f = make_flow(ctx, t, nu=nu, cfl=0.5, inlet_speed=U, inlet_profile=1, face=faces, convection=1,
                     poisson_tol=1e-10)
    try:
        c = t.centers()
        s = c[:, 1] / H
        exact = 6.0 * U * s * (1.0 - s)
        n = t.n
        f.set_state(exact, np.zeros(n), np.zeros(n))
        for _ in range(100):
            r = f.step()
        u, v, w, p = f.get_state()
        core = (c[:, 0] > 0.3) & (c[:, 0] < 0.7)
        err = np.linalg.norm(u[core] - exact[core]) / np.linalg.norm(exact[core])
        print(f"Poiseuille: rel L2 {err:.3e}, |v| {np.abs(v).max():.2e}, t {r.t:.3f}", flush=True)
        assert err < 0.02
        assert np.abs(v).max() < 0.02 * U and np.abs(w).max() < 1e-12
        # pressure gradient of the core, by least squares in x: every stage is
        # projected with its own step, so p is the pressure whose gradient
        # -12 nu U / H^2 balances the parabola's viscous force
        mid = (c[:, 0] > 0.25) & (c[:, 0] < 0.75) & (c[:, 1] > 0.2) & (c[:, 1] < 0.8)
        slope = np.polyfit(c[mid, 0], p[mid], 1)[0]
        expect = -12.0 * nu * U / H ** 2
        print(f"dp/dx {slope:.5f} against {expect:.5f}", flush=True)
        assert abs(slope - expect) <= 0.02 * abs(expect)
    finally:
        f.close()
        d.close()


def test_solid_leaves_and_determinism(ctx, orc):
    t = jump_tree(orc)
    c = t.centers()
    solid = (np.abs(c[:, 0] - 0.9) < 0.15) & (np.abs(c[:, 1] - 0.5) < 0.15) & (np.abs(c[:, 2] - 0.5) < 0.15)
    for i in np.nonzero(solid)[0]:
        t.set_mask(int(i), 2)
    faces = (nbl.OCT_INLET, nbl.OCT_OUTLET, nbl.OCT_WALL, nbl.OCT_WALL, nbl.OCT_SLIP, nbl.OCT_SLIP)
    out = []
    for _ in range(2):
        d, f = make_flow(ctx, t, nu=0.01, cfl=0.5, face=faces, convection=1, poisson_tol=1e-10)
        try:
            n = t.n
            f.set_state(np.ones(n), np.zeros(n), np.zeros(n))
            reps = [f.step().as_dict() for _ in range(4)]
            st = f.get_state()
            out.append((reps, st))
            assert all(np.all(a[solid] == 0.0) for a in st[:3])
            assert reps[-1]["divergence_max"] < 1e-7
            assert abs(reps[-1]["inflow"] - reps[-1]["outflow"]) < 1e-12
        finally:
            f.close()
            d.close()
    (ra, sa), (rb, sb) = out
    assert [{k: v for k, v in r.items() if k != "solve_ms"} for r in ra] == \
           [{k: v for k, v in r.items() if k != "solve_ms"} for r in rb]
    for a, b in zip(sa, sb):
        assert np.array_equal(a, b)


def test_flow_options_validated(ctx, orc):
    t = jump_tree(orc)
    k, l, m = t.dump()
    d = nbl.Octree(ctx, k, l, m, t.origin, t.extent, t.max_level)
    try:
        for kw in (dict(cfl=0.0), dict(convection=2), dict(inlet_profile=3), dict(face=(7, 0, 0, 0, 0, 0)),
                   dict(poisson_max_iters=0), dict(precision=5)):
            with pytest.raises(nbl.NblError):
                nbl.OctreeFlow(d, nbl.oct_flow_cfg(**kw))
    finally:
        d.close()


def duct_runs(ctx, orc, n, steps, dt, nu=0.05):
    """Developing flow in a 2 x 1 x 1 duct (uniform inflow, outlet, no-slip
    side walls) on n^3 cells of (2/n, 1/n, 1/n): the octree step on a
    uniformly refined tree and the MAC step on the same cells, both from
    u = 1, WENO-5, the same fixed dt.  Returns both cell-centre velocities
    in the grid's cell order and the cells' x centres."""
    L = int(round(np.log2(n)))
    t = orc.octree((0.0, 0.0, 0.0), (2.0, 1.0, 1.0), L)
    t.refine_uniform(L)
    walls = (nbl.OCT_WALL,) * 4
    d, f = make_flow(ctx, t, nu=nu, inlet_speed=1.0, face=(nbl.OCT_INLET, nbl.OCT_OUTLET) + walls, convection=1,
                     poisson_tol=1e-10)
    try:
        f.set_state(np.ones(t.n), np.zeros(t.n), np.zeros(t.n))
        for _ in range(steps):
            f.step(dt)
        uo = f.get_state()[:3]
    finally:
        f.close()
        d.close()
    c = t.centers()
    ix = np.floor(c[:, 0] * n / 2.0).astype(np.int64)
    iy = np.floor(c[:, 1] * n).astype(np.int64)
    iz = np.floor(c[:, 2] * n).astype(np.int64)
    cell = ix + n * (iy + n * iz)
    assert np.array_equal(np.sort(cell), np.arange(n ** 3))
    faces = [dict(type=nbl.BC_INLET, profile=nbl.INLET_UNIFORM, inlet_speed=1.0, inlet_axis=0),
             dict(type=nbl.BC_OUTLET)] + [dict(type=nbl.BC_WALL) for _ in range(4)]
    g = nbl.Grid(n, n, n, 2.0 / n, 1.0 / n, 1.0 / n, faces=faces)
    cfg = nbl.flow_cfg(reynolds=1.0 / nu, ref_length=1.0, velocity_scale=1.0, convection=nbl.CONV_WENO5,
                       poisson_tol=1e-10)
    run = nbl.Run(ctx, g, cfg)
    try:
        run.initialize()
        for _ in range(steps):
            run.step(dt)
        um = [run.cell_velocity(q) for q in range(3)]
    finally:
        run.close()
    o = [np.empty(n ** 3) for _ in range(3)]
    for q in range(3):
        o[q][cell] = uo[q]
    xc = (np.arange(n ** 3) % n + 0.5) * 2.0 / n
    return o, um, xc


def test_agrees_with_mac_step_on_a_uniform_tree(ctx, orc):
    """On a uniformly refined tree the collocated octree step and the MAC step
    discretize the same developing duct flow: their cell velocities differ by
    the discretizations' truncation error, which shrinks under refinement."""
    errs = []
    for n in (16, 32):
        o, m, xc = duct_runs(ctx, orc, n, steps=100, dt=0.0025)
        core = (xc > 0.25) & (xc < 1.25)  # away from the outlet's different closure
        num = np.sqrt(sum(np.sum((a[core] - b[core]) ** 2) for a, b in zip(o, m)))
        den = np.sqrt(sum(np.sum(b[core] ** 2) for b in m))
        errs.append(num / den)
        print(f"duct {n}^3: octree vs MAC rel L2 {errs[-1]:.3e}", flush=True)
    assert errs[0] < 0.1
    assert errs[1] < 0.7 * errs[0]
