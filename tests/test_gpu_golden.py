"""The HIP path against the golden fixtures written by the reference compiled
from its own sources (tests/golden/make_golden.py), through the C-ABI.

Per-element work is bit-exact to the reference; solves and steps contain
reductions in a different (tree) order, so fields are compared with relative
tolerances stated per test, iteration counts exactly (or within one where a
residual sits at the tolerance), and integral histories (kinetic energy,
enstrophy, drag and lift) over the whole fixture.
"""
from pathlib import Path

import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    p = GOLD / name
    if not p.exists():
        pytest.skip(f"fixture {name} not generated")
    return np.load(p, allow_pickle=False)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    s = np.linalg.norm(b)
    return np.linalg.norm(a - b) / s if s > 0 else np.linalg.norm(a - b)


def rep_array(reps, key):
    return np.array([getattr(r, key) for r in reps])


def test_mg_block_solution(ctx):
    """core/tests/test_flow.cpp:63-117's operator: inlet/outlet/walls/periodic
    with a solid block, MG-PCG and Jacobi-CG to 1e-10."""
    f = load("mg_block.npz")
    faces = [dict(type=nbl.BC_INLET), dict(type=nbl.BC_OUTLET), dict(type=nbl.BC_WALL),
             dict(type=nbl.BC_WALL), dict(type=nbl.BC_PERIODIC), dict(type=nbl.BC_PERIODIC)]
    solid = np.zeros((16, 16, 32), np.uint8)
    solid[6:10, 0:6, 12:18] = 1
    g = nbl.Grid(32, 16, 16, 1.0 / 16, faces=faces, solid=solid)
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-10, max_iters=100)
    try:
        x, rep = ps.solve(f["b"])
    finally:
        ps.close()
    assert rep.converged and rep.iterations == int(f["iterations"])
    assert rep.residual == pytest.approx(float(f["residual"]), rel=1e-3)
    assert rel(x, f["x"]) < 1e-9
    pc = nbl.Poisson(ctx, g, kind=1, tol=1e-10, max_iters=5000)
    try:
        xc, repc = pc.solve(f["b"])
    finally:
        pc.close()
    assert abs(repc.iterations - int(f["cg_iterations"])) <= 1
    assert rel(xc, f["x_cg"]) < 1e-8


def test_cube48_twenty_steps(ctx):
    f = load("cube48_20.npz")
    g, cfg = cases.cube_channel(nx=48, ny=16, nz=16, res=4)
    run = nbl.Run(ctx, g, cfg)
    try:
        run.set_state(*cases.cube_channel_state(g))
        reps = [run.step() for _ in range(20)]
        u, v, w, p, *_ = run.get_state()
    finally:
        run.close()
    assert np.array_equal(rep_array(reps, "poisson_iters"), f["rep_poisson_iters"])
    assert np.allclose(rep_array(reps, "dt"), f["rep_dt"], rtol=1e-12, atol=0)
    assert np.allclose(rep_array(reps, "cd"), f["rep_cd"], rtol=1e-9, atol=1e-12)
    for a, name in zip((u, v, w, p), "uvwp"):
        assert rel(a, f[name]) < 1e-10, name


def test_tgv64_hundred_steps(ctx):
    """64^3 periodic Taylor-Green vortex, 100 steps: kinetic energy and
    enstrophy histories, step reports, and 4096 sampled values per field."""
    f = load("tgv64_100.npz")
    g, cfg = cases.taylor_green(64)
    run = nbl.Run(ctx, g, cfg)
    ke, ens, reps = [], [], []
    try:
        run.set_state(*cases.taylor_green_state(g))
        for _ in range(100):
            reps.append(run.step())
            u, v, w, *_ = run.get_state()
            a, b = cases.kinetic_energy_enstrophy(g, u, v, w)
            ke.append(a)
            ens.append(b)
        u, v, w, p, *_ = run.get_state()
    finally:
        run.close()
    assert np.allclose(ke, f["ke"], rtol=1e-12, atol=0)
    assert np.allclose(ens, f["enstrophy"], rtol=1e-10, atol=0)
    iters = rep_array(reps, "poisson_iters")
    assert np.abs(iters - f["rep_poisson_iters"]).max() <= 1
    assert np.allclose(rep_array(reps, "t"), f["rep_t"], rtol=1e-12, atol=0)
    assert (rep_array(reps, "poisson_residual") <= 1e-8).all()
    for a, name in zip((u, v, w, p), "uvwp"):
        assert np.linalg.norm(a) == pytest.approx(float(f[f"{name}_l2"]), rel=1e-10)
        d = np.abs(a[f[f"{name}_idx"]] - f[f"{name}_val"]).max()
        assert d <= 1e-10 * max(float(f[f"{name}_max"]), 1e-30), name


def test_cube192_history(ctx):
    """The 192x64x64 wall-mounted-cube channel (Re_h 500, WENO-5, MG-PCG 1e-8)
    from the seeded perturbed state: per-step dt, drag and lift coefficients and
    Poisson iteration counts over the fixture's whole history."""
    f = load("cube192_history.npz")
    n = len(f["rep_step"])
    g, cfg = cases.cube_channel()
    run = nbl.Run(ctx, g, cfg)
    try:
        run.set_state(*cases.cube_channel_state(g))
        reps = [run.step() for _ in range(n)]
        st = run.get_state() if "u_l2" in f.files else None
    finally:
        run.close()
    iters = rep_array(reps, "poisson_iters")
    assert np.abs(iters - f["rep_poisson_iters"]).max() <= 1
    assert np.allclose(rep_array(reps, "dt"), f["rep_dt"], rtol=1e-10, atol=0)
    assert np.allclose(rep_array(reps, "cd"), f["rep_cd"], rtol=1e-8, atol=1e-12)
    assert np.allclose(rep_array(reps, "cl"), f["rep_cl"], rtol=1e-6, atol=1e-9)
    assert (rep_array(reps, "poisson_residual") <= 1e-8).all()
    assert np.array_equal(rep_array(reps, "accepted"), f["rep_accepted"])
    if st is not None:
        for a, name in zip(st[:4], "uvwp"):
            assert np.linalg.norm(a) == pytest.approx(float(f[f"{name}_l2"]), rel=1e-9)
            got = a[f[f"{name}_idx"]]
            assert np.abs(got - f[f"{name}_val"]).max() <= 1e-9 * np.abs(f[f"{name}_val"]).max(), name


# ---- the mixed form (FP32 V-cycle inside the FP64 flexible CG) ------------------
# Tolerances of the mixed form against the reference's FP64 fixtures: the FP64
# outer residual reaches the same 1e-8 at every solve, drag and lift within 1e-6
# relative (1e-9 absolute where they cross zero), and the integral histories
# and fields within the rounding the FP32 preconditioner leaves inside that
# tolerance.
@pytest.mark.parametrize("prec", [nbl.PREC_MIXED, nbl.PREC_MIXED_BF16], ids=["fp32", "bf16"])
def test_tgv64_hundred_steps_mixed(ctx, prec):
    f = load("tgv64_100.npz")
    g, cfg = cases.taylor_green(64)
    run = nbl.Run(ctx, g, cfg, nbl.solver_opts(prec))
    ke, ens, reps = [], [], []
    try:
        run.set_state(*cases.taylor_green_state(g))
        for _ in range(100):
            reps.append(run.step())
            u, v, w, *_ = run.get_state()
            a, b = cases.kinetic_energy_enstrophy(g, u, v, w)
            ke.append(a)
            ens.append(b)
        u, v, w, p, *_ = run.get_state()
    finally:
        run.close()
    ke_err = np.abs(np.array(ke) / f["ke"] - 1).max()
    ens_err = np.abs(np.array(ens) / f["enstrophy"] - 1).max()
    iters = rep_array(reps, "poisson_iters")
    print(f"TGV64 mixed ({prec}): KE {ke_err:.2e}, enstrophy {ens_err:.2e}, iterations "
          f"{iters.sum()} vs {int(f['rep_poisson_iters'].sum())}", flush=True)
    assert ke_err <= 1e-10 and ens_err <= 1e-9
    assert np.abs(iters - f["rep_poisson_iters"]).max() <= 3
    assert (rep_array(reps, "poisson_residual") <= 1e-8).all()
    for a, name in zip((u, v, w, p), "uvwp"):
        e = abs(np.linalg.norm(a) / float(f[f"{name}_l2"]) - 1)
        d = np.abs(a[f[f"{name}_idx"]] - f[f"{name}_val"]).max() / max(float(f[f"{name}_max"]), 1e-30)
        print(f"  {name}: l2 {e:.2e}, sampled max {d:.2e}", flush=True)
        assert e <= 1e-8 and d <= 1e-7, name


@pytest.mark.parametrize("prec", [nbl.PREC_MIXED, nbl.PREC_MIXED_BF16], ids=["fp32", "bf16"])
def test_cube192_history_mixed(ctx, prec):
    """The 192x64x64 cube channel over the fixture's 500 steps with the FP32
    V-cycle (and its BF16 smoother-scratch form): residual <= 1e-8 at every
    solve, drag and lift within 1e-6."""
    f = load("cube192_history.npz")
    n = len(f["rep_step"])
    g, cfg = cases.cube_channel()
    run = nbl.Run(ctx, g, cfg, nbl.solver_opts(prec))
    try:
        run.set_state(*cases.cube_channel_state(g))
        reps = [run.step() for _ in range(n)]
        st = run.get_state()
    finally:
        run.close()
    cd, cl = rep_array(reps, "cd"), rep_array(reps, "cl")
    cd_err = np.abs(cd / f["rep_cd"] - 1).max()
    cl_err = (np.abs(cl - f["rep_cl"]) / np.maximum(np.abs(f["rep_cl"]), 1e-3)).max()
    iters = rep_array(reps, "poisson_iters")
    print(f"cube192 mixed ({prec}), {n} steps: cd {cd_err:.2e}, cl {cl_err:.2e}, iterations {iters.sum()} vs "
          f"{int(f['rep_poisson_iters'].sum())}, max residual {rep_array(reps, 'poisson_residual').max():.2e}",
          flush=True)
    assert (rep_array(reps, "poisson_residual") <= 1e-8).all()
    assert np.abs(iters - f["rep_poisson_iters"]).max() <= 3
    assert np.allclose(rep_array(reps, "dt"), f["rep_dt"], rtol=1e-8, atol=0)
    assert np.allclose(cd, f["rep_cd"], rtol=1e-6, atol=1e-9)
    assert np.allclose(cl, f["rep_cl"], rtol=1e-6, atol=1e-9)
    assert np.array_equal(rep_array(reps, "accepted"), f["rep_accepted"])
    for a, name in zip(st[:4], "uvwp"):
        e = abs(np.linalg.norm(a) / float(f[f"{name}_l2"]) - 1)
        d = np.abs(a[f[f"{name}_idx"]] - f[f"{name}_val"]).max() / np.abs(f[f"{name}_val"]).max()
        print(f"  {name}: l2 {e:.2e}, sampled max {d:.2e}", flush=True)
        assert e <= 1e-8 and d <= 1e-6, name
