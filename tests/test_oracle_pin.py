"""Pin the oracle restatement against the reference compiled from its own
sources (oracle/_ref): every result must be bit-identical, because both run
the same serial FP64 arithmetic in the same order.

Skipped where oracle/_ref is not built (it is only built where the reference
tree is mounted); the committed golden fixtures cover the same ground there
(test_oracle_golden.py).
"""
import numpy as np
import pytest

import cases
import nbl


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def grids():
    """Small grids exercising every boundary kind, solids and 2-D."""
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
    return out


@pytest.mark.parametrize("name,g", grids(), ids=[n for n, _ in grids()])
def test_apply_and_diagonal_bitwise(orc, ref, name, g):
    x = cases.normal(7, g.cells)
    assert same(orc.apply(g, x), ref.apply(g, x))
    assert same(orc.diag(g), ref.diag(g))


@pytest.mark.parametrize("name,g", grids(), ids=[n for n, _ in grids()])
def test_cg_and_mg_bitwise(orc, ref, name, g):
    b = cases.poisson_rhs(g, 11)
    xo, ro = orc.cg(g, b, 1e-10, 3000)
    xr, rr = ref.cg(g, b, 1e-10, 3000)
    assert same(xo, xr) and (ro.iterations, ro.residual) == (rr.iterations, rr.residual)
    xo, ro = orc.mg(g, b, 1e-10, 100)
    xr, rr = ref.mg(g, b, 1e-10, 100)
    assert same(xo, xr)
    assert (ro.iterations, ro.residual, ro.converged) == (rr.iterations, rr.residual, rr.converged)


def _ns_pair(orc, ref, g, cfg, state=None):
    a, b = orc.ns(g, cfg), ref.ns(g, cfg)
    for s in (a, b):
        s.initialize()
        if state is not None:
            s.set_state(*state)
    return a, b


def _check_steps(a, b, n, dt=0.0):
    for _ in range(n):
        ra, rb = a.step(dt), b.step(dt)
        assert ra.as_dict() == rb.as_dict()
    for x, y in zip(a.get_state(), b.get_state()):
        assert same(x, y)


@pytest.mark.parametrize("conv", [nbl.CONV_CENTRAL, nbl.CONV_WENO5])
@pytest.mark.parametrize("name,g", grids(), ids=[n for n, _ in grids()])
def test_ns_steps_bitwise(orc, ref, name, g, conv):
    cfg = nbl.flow_cfg(reynolds=200.0, convection=conv, poisson_tol=1e-9, cfl=0.5,
                       body_force=(0.01, 0.0, 0.0))
    u, v, w, p = (cases.normal(s, n) * 0.3 for s, n in zip((1, 2, 3, 4), g.shapes))
    a, b = _ns_pair(orc, ref, g, cfg, (u, v, w, p))
    assert a.compute_dt() == b.compute_dt()
    _check_steps(a, b, 3)


def test_tgv16_steps_bitwise(orc, ref):
    g, cfg = cases.taylor_green(16)
    a, b = _ns_pair(orc, ref, g, cfg, cases.taylor_green_state(g))
    _check_steps(a, b, 4)


def test_cube_small_steps_bitwise(orc, ref):
    g, cfg = cases.cube_channel(nx=48, ny=16, nz=16, res=4)
    st = cases.cube_channel_state(g)
    a, b = _ns_pair(orc, ref, g, cfg, st)
    _check_steps(a, b, 3)
    ra = a.step()
    assert ra.cd != 0.0


def test_morton_bitwise(orc, ref):
    rng = np.random.default_rng(5)
    xyz = rng.integers(0, 1 << 21, size=(2000, 3), dtype=np.uint32)
    for x, y, z in xyz:
        assert orc.morton_encode(x, y, z) == ref.morton_encode(x, y, z)


def _oct_pair(orc, ref, origin, extent, ml):
    return orc.octree(origin, extent, ml), ref.octree(origin, extent, ml)


def _check_oct(a, b):
    for x, y in zip(a.dump(), b.dump()):
        assert same(x, y)
    assert a.balanced() == b.balanced()
    for i in range(a.n):
        for d in range(6):
            assert a.neighbors(i, d) == b.neighbors(i, d)
    assert same(a.centers(), b.centers())
    x = cases.normal(3, a.n)
    assert same(a.apply(x), b.apply(x))
    assert same(a.diagonal(), b.diagonal())
    # octreeTransportStep (adaptive/transport.cpp:24-60): the reference test's
    # fields (test_adaptive.cpp:30-34) and seeded random velocities
    c = a.centers()
    phi = 1.0 + np.sin(6.0 * c[:, 0]) * np.cos(5.0 * c[:, 1])
    for u, v, w in ((np.full(a.n, 0.3), np.full(a.n, 0.1), np.zeros(a.n)),
                    (cases.normal(41, a.n), cases.normal(42, a.n), cases.normal(43, a.n))):
        pa, ia = a.transport(u, v, w, phi, 0.02)
        pb, ib = b.transport(u, v, w, phi, 0.02)
        assert same(pa, pb) and ia == ib
    # PhysicalSensorEvaluator::evaluate (adaptive_core.cpp:73-176): seeded
    # fields over three evaluations, so the carried speed / residual state and
    # the stagnation flag take both values
    f = [cases.normal(51 + k, a.n) for k in range(4)]
    scale = 1.0 + np.abs(cases.normal(55, a.n))
    spa = spb = rpa = rpb = np.zeros(a.n)
    for _ in range(3):
        sa, spa, rpa = a.sensors(*f, 1.3e-3, spa, rpa)
        sb, spb, rpb = b.sensors(*f, 1.3e-3, spb, rpb)
        for k in sa:
            assert same(sa[k], sb[k]), k
        assert same(spa, spb) and same(rpa, rpb)
        f = [f[0] * scale, f[1] * scale, f[2] * scale, f[3]]
    bb = a.apply(x)
    xa, ra = a.solve(bb, 1e-10, 3000)
    xb, rb = b.solve(bb, 1e-10, 3000)
    assert same(xa, xb) and (ra.iterations, ra.residual) == (rb.iterations, rb.residual)


def test_octree_box_refine_bitwise(orc, ref):
    a, b = _oct_pair(orc, ref, (0, 0, 0), (1, 1, 1), 5)
    for t in (a, b):
        t.refine_uniform(2)
        t.refine_box((0, 0, 0), (0.5, 0.5, 0.5), 1)
    _check_oct(a, b)


def test_octree_refine_storm_bitwise(orc, ref):
    a, b = _oct_pair(orc, ref, (-7.0, -1.5, -3.2), (14.0, 3.0, 6.4), 7)
    a.refine_uniform(2)
    b.refine_uniform(2)
    rng = np.random.default_rng(12345)
    for _ in range(60):
        i = int(rng.integers(0, a.n))
        k, l, _m = a.dump()
        if l[i] < 7:
            a.refine_key(k[i])
            b.refine_key(k[i])
    for i in range(0, a.n, 7):
        a.set_mask(i, 2)
        b.set_mask(i, 2)
    _check_oct(a, b)
    assert a.balanced()
