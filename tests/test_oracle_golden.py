"""Pin the oracle restatement against the golden fixtures written by the
reference compiled from its own sources (tests/golden/make_golden.py).
Bit-exact: the oracle runs the same serial FP64 arithmetic as the reference."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import cases
import nbl

GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    p = GOLD / name
    if not p.exists():
        pytest.skip(f"fixture {name} not generated")
    return np.load(p, allow_pickle=False)


def check_reports(reps, f):
    for k, _ in nbl.StepReport._fields_:
        got = np.array([getattr(r, k) for r in reps])
        assert np.array_equal(got, f[f"rep_{k}"][: len(reps)]), k


def test_tgv16_bitwise(orc):
    f = load("tgv16.npz")
    g, cfg = cases.taylor_green(16)
    s = orc.ns(g, cfg)
    s.set_state(*cases.taylor_green_state(g))
    reps = []
    for q in range(10):
        reps.append(s.step())
        if q in (0, 9):
            u, v, w, p, *_ = s.get_state()
            for name, a in zip("uvwp", (u, v, w, p)):
                assert np.array_equal(a, f[f"{name}{q + 1}"])
    check_reports(reps, f)


def test_mg_block_bitwise(orc):
    f = load("mg_block.npz")
    faces = [dict(type=nbl.BC_INLET), dict(type=nbl.BC_OUTLET), dict(type=nbl.BC_WALL),
             dict(type=nbl.BC_WALL), dict(type=nbl.BC_PERIODIC), dict(type=nbl.BC_PERIODIC)]
    solid = np.zeros((16, 16, 32), np.uint8)
    solid[6:10, 0:6, 12:18] = 1
    g = nbl.Grid(32, 16, 16, 1.0 / 16, faces=faces, solid=solid)
    x, rep = orc.mg(g, f["b"], 1e-10, 100)
    assert np.array_equal(x, f["x"])
    assert (rep.iterations, rep.residual) == (int(f["iterations"]), float(f["residual"]))
    assert rep.iterations <= 50 and rep.residual < 1e-8  # core/tests/test_flow.cpp:107-109
    xc, repc = orc.cg(g, f["b"], 1e-10, 5000)
    assert np.array_equal(xc, f["x_cg"]) and repc.iterations == int(f["cg_iterations"])


def test_cube48_twenty_steps_bitwise(orc):
    f = load("cube48_20.npz")
    g, cfg = cases.cube_channel(nx=48, ny=16, nz=16, res=4)
    s = orc.ns(g, cfg)
    s.set_state(*cases.cube_channel_state(g))
    reps = [s.step() for _ in range(20)]
    check_reports(reps, f)
    for name, a in zip("uvwp", s.get_state()[:4]):
        assert np.array_equal(a, f[name])


def _oct_storm(orc):
    t = orc.octree((-7.0, -1.5, -3.2), (14.0, 3.0, 6.4), 7)
    t.refine_uniform(2)
    picks = cases.splitmix64(4242, 80)
    for q in range(80):
        k, l, _m = t.dump()
        i = int(picks[q] % np.uint64(t.n))
        if l[i] < 7:
            t.refine_key(k[i])
    for i in range(0, t.n, 5):
        t.set_mask(i, 2)
    return t


def _nbrs(t):
    nb = np.full((t.n, 6, 4), -1, np.int64)
    for i in range(t.n):
        for d in range(6):
            for q, j in enumerate(t.neighbors(i, d)):
                nb[i, d, q] = j
    return nb


def test_octree_fixtures_bitwise(orc):
    f = load("octree_small.npz")
    t = orc.octree((0, 0, 0), (1, 1, 1), 5)
    t.refine_uniform(2)
    t.refine_box((0, 0, 0), (0.5, 0.5, 0.5), 1)
    k, l, m = t.dump()
    assert np.array_equal(k, f["jump_keys"]) and np.array_equal(l, f["jump_levels"])
    assert np.array_equal(_nbrs(t), f["jump_nbrs"])
    assert np.array_equal(t.apply(f["jump_x"]), f["jump_Ax"])
    assert np.array_equal(t.diagonal(), f["jump_diag"])
    assert np.array_equal(t.centers(), f["jump_centers"])
    xs, rep = t.solve(f["jump_Ax"], 1e-10, 3000)
    assert np.array_equal(xs, f["jump_sol"]) and rep.iterations == int(f["jump_iters"])
    t = _oct_storm(orc)
    k, l, m = t.dump()
    assert np.array_equal(k, f["storm_keys"]) and np.array_equal(m, f["storm_masks"])
    assert np.array_equal(_nbrs(t), f["storm_nbrs"])
    assert np.array_equal(t.apply(f["storm_x"]), f["storm_Ax"])
    assert np.array_equal(t.diagonal(), f["storm_diag"])
    assert t.balanced() and bool(f["storm_balanced"])


def build_c3(orc, max_level, base):
    t = orc.octree(cases.C3_ORIGIN, cases.C3_EXTENT, max_level)
    t.refine_uniform(base)
    while True:
        k, l, _m = t.dump()
        sel = cases.c3_predicate(t.centers()) & (l < max_level)
        if not sel.any():
            return t
        t.refine_flagged_pass(k[sel], l[sel])


def tree_hash(k, l):
    o = np.argsort(k)
    return hashlib.sha256(np.ascontiguousarray(k[o]).tobytes() + np.ascontiguousarray(l[o]).tobytes()).hexdigest()


def test_c3_small_tree_keyset(orc):
    f = load("c3_tree_L6.npz")
    t = build_c3(orc, 6, 3)
    k, l, _m = t.dump()
    assert t.n == int(f["n"])
    assert tree_hash(k, l) == str(f["sha256"])
    assert np.array_equal(np.bincount(l, minlength=7), f["level_hist"])
    assert t.balanced()


def test_tgv64_history_head(orc):
    """First 3 of the 100 reference steps of the 64^3 Taylor-Green case."""
    f = load("tgv64_100.npz")
    g, cfg = cases.taylor_green(64)
    s = orc.ns(g, cfg)
    s.set_state(*cases.taylor_green_state(g))
    reps = []
    for q in range(3):
        reps.append(s.step())
        u, v, w, *_ = s.get_state()
        ke, ens = cases.kinetic_energy_enstrophy(g, u, v, w)
        assert ke == f["ke"][q] and ens == f["enstrophy"][q]
    check_reports(reps, f)
