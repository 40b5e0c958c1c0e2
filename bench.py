#!/usr/bin/env python3
"""Benchmark of the nabla_solve per-step hot path on one MI355X.

Workload (one "step" = one SSP-RK3 MAC time step, three projections):
192x64x64 cube channel (h = 1, 16 cells per h), uniform inlet U = 1 with a
seeded 1 % perturbation, no-slip floor/ceiling, periodic span, convective
outlet, staircase cube [3,4]x[0,1]x[1.5,2.5], Re_h 500, CFL 0.7, WENO-5,
MG-PCG pressure solve to 1e-8 relative residual; FP64 throughout (the plain
path).  The same steps are also run with the FP32 V-cycle inside the FP64
flexible CG ("mixed") and reported beside it.

Prints ONE JSON line (rank 0).  `value` is whole-job Mcell-steps/s of the FP64
path with the state resident in HBM.  `roofline` is the FP64 damped-Jacobi
sweep (the workload's dominant kernel) on a 256^3 level of the same system
class (walls and one outlet face: non-singular, as the channel's), whose
436 MB per launch stream from HBM, timed with HIP events on the solver's
stream over 100 back-to-back launches; the channel's own 192x64x64 level
(20 MB per launch, resident in the 256 MB Infinity Cache) is reported beside
it as `roofline_cube192_infinity_cache`.  `cpu_baseline` is the reference
solver (oracle/_ref, else the oracle restatement) on one host core over a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "nabla-ai_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import numpy as np  # noqa: E402

import cases  # noqa: E402
import nbl  # noqa: E402

METRIC = "Mcell-steps/s + MG-PCG solve ms at rel. resid 1e-8; p50/p99 served-step latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_steps(ctx, g, cfg, st, precision, warmup, steps, timing=None, loops=False):
    run = nbl.Run(ctx, g, cfg, nbl.solver_opts(precision=precision, loops=loops))
    run.set_state(*st)
    for _ in range(warmup):
        run.step()
    reps, solve_ms, iters = [], [], []
    if timing is not None:
        run.sweep_timing(1)
    t0 = time.perf_counter()
    for s in range(steps):
        r = run.step()
        ms, it, _vc = run.last_solves()
        reps.append(r.as_dict())
        solve_ms.extend(ms)
        iters.extend(it)
        if (s + 1) % max(1, steps // 5) == 0:
            log(f"prec={precision} step {s + 1}/{steps}")
    elapsed = time.perf_counter() - t0
    if timing is not None:
        timing.update(zip(("avg_ms", "launches", "bytes"), run.sweep_timing(0)))
    state = run.get_state()
    run.close()
    return reps, solve_ms, iters, elapsed, state


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def grid256(singular=False):
    """256^3 unit box: all walls (singular) or walls with an outlet on the +z
    face (non-singular, the channel's system class)."""
    faces = [dict(type=nbl.BC_WALL)] * 6 if singular else \
        [dict(type=nbl.BC_WALL)] * 5 + [dict(type=nbl.BC_OUTLET)]
    return nbl.Grid(256, 256, 256, 1.0 / 256, faces=faces)


def roofline(ctx, g, reps=100, precision=nbl.PREC_FP64, kernel=0, key=None):
    """One solve kernel in a loop of back-to-back launches on the solver's
    stream between two HIP events: 0 the finest-level sweep (FP64, or FP32 as
    in the mixed V-cycle), 1 the FP64 operator application with p.Ap (the
    outer CG's SpMV), 4 the finest-level down kernel of the
    two-kernels-per-level cycle."""
    vc = nbl.VCYCLE_TILED if kernel == 4 else nbl.VCYCLE_AUTO
    ps = nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=5000, opts=nbl.solver_opts(precision=precision, vcycle=vc))
    ps.profile(kernel, 10)
    ms, by = ps.profile(kernel, reps)
    ps.close()
    ach = by / (ms * 1e-3) / 1e9
    fp32 = precision != nbl.PREC_FP64
    t = "float" if fp32 else "double"
    name = {1: "k_fcg_apply<2> (A p and p.Ap, FP64; 18 B/cell)",
            4: f"k_down<{t}, DS_B> (both pre-sweeps, residual and restriction in LDS tiles; "
               f"{2 * (4 if fp32 else 8) + 2} B/cell + the coarse residual)"}.get(
        kernel, f"k_sweep<{t}, SW_PLAIN> (finest-level damped-Jacobi sweep; {14 if fp32 else 26} B/cell)")
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(key) if key else None,
            "kernel": name, "avg_launch_us": round(ms * 1e3, 3), "bytes_per_launch": by, "cells": g.cells,
            "grid": f"{g.nx}x{g.ny}x{g.nz}", "launches_timed": reps,
            "measured": "HIP events on the solver's stream around back-to-back launches"}


def pmc_traffic(key):
    """HBM bytes per launch from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes, copy-calibrated), or None."""
    p = ROOT / "profiles" / "k_sweep_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text()).get(key)
        return None if d is None else d.get("bytes_per_launch")
    except (ValueError, OSError):
        return None


def roofline_insitu(sweep, g):
    """Finest-level plain sweep timed inside the timed steps: every block of a
    launch stores the device wall clock at its start and when all its waves
    are done; a launch lasts from its first block start to its last block end
    (launches a converged solve skipped are left out).  This leaves out the
    dispatch cost every launch pays; the 20 MB the sweep touches on this level
    sit in the 256 MB Infinity Cache, so the figure is not an HBM one."""
    ms, n, by = sweep["avg_ms"], sweep["launches"], sweep["bytes"]
    ach = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic("cube192x64x64"),
            "kernel": "k_sweep<double, SW_PLAIN> on the 192x64x64 level (pre-smoothing sweep 2 of every V-cycle)",
            "avg_launch_us": round(ms * 1e3, 3), "launches_timed": n, "bytes_per_launch": by, "cells": g.cells,
            "measured": "device wall-clock stamps per block (first block start to last block end), every launch "
                        "inside the timed steps"}


def coarse_solve(ctx, g, reps=50):
    """Average time of one coarsest-level solve (24x8x8 here) for both coarse
    modes: the direct inverse (default) and the reference's CG(1e-6, 500)."""
    out = {}
    for name, mode in (("direct", nbl.COARSE_DIRECT), ("cg", nbl.COARSE_CG)):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=5000, opts=nbl.solver_opts(coarse=mode))
        ms, _by = ps.profile(3, reps)
        ps.close()
        out[f"{name}_us"] = round(ms * 1e3, 2)
    return out


def octree_flow(d, masks, steps, warmup=3, precision=nbl.PREC_MIXED, seed=20240601):
    """Flow steps on the cube-wake octree's leaves (nbl_octree_flow_step):
    Re_h 500 (nu = 1/500, h = U = 1), CFL 0.7, WENO-5; uniform inflow at x = 0,
    outlet at x = 12, no-slip floor and ceiling, free-slip side planes (the
    uniform-grid cube channel is periodic in the span; the octree has no
    periodic faces).  Initial u = 1 + 0.01 xi on fluid leaves (xi ~ U(-1,1),
    splitmix64 seed), v = w = p = 0.  Each step: three RK stages, each followed
    by a projection through the octree multigrid-preconditioned CG (FP32
    V-cycle by default) to 1e-8, warm-started from the last pressure."""
    n = masks.size
    fl = masks == 0
    u = np.where(fl, 1.0 + 0.01 * (2.0 * cases.uniform01(seed, n) - 1.0), 0.0)
    z = np.zeros(n)
    cfg = nbl.oct_flow_cfg(nu=1.0 / 500.0, cfl=0.7, inlet_speed=1.0, inlet_profile=0,
                           face=(nbl.OCT_INLET, nbl.OCT_OUTLET, nbl.OCT_WALL, nbl.OCT_WALL,
                                 nbl.OCT_SLIP, nbl.OCT_SLIP),
                           convection=1, poisson_tol=1e-8, poisson_max_iters=500, precision=precision)
    f = nbl.OctreeFlow(d, cfg)
    try:
        f.set_state(u, z, z)
        for _ in range(warmup):
            f.step()
        reps = []
        t0 = time.perf_counter()
        for _ in range(steps):
            reps.append(f.step())
        el = time.perf_counter() - t0
        _u, _v, _w, p = f.get_state()
    finally:
        f.close()
    it = np.array([r.poisson_iters for r in reps])
    solve = float(sum(r.solve_ms for r in reps))
    last = reps[-1]
    return {"leaves": int(n), "fluid_leaves": int(fl.sum()), "steps": steps, "warmup": warmup,
            "precision": {nbl.PREC_FP64: "fp64", nbl.PREC_MIXED: "fp32_vcycle",
                          nbl.PREC_MIXED_BF16: "bf16_scratch"}[precision],
            "ms_per_step": round(el / steps * 1e3, 3),
            "mleaf_steps_per_s": round(n * steps / el / 1e6, 2),
            "projection_ms_per_step": round(solve / steps, 3),
            "projection_share": round(solve * 1e-3 / el, 3),
            "poisson_iters_per_step_mean": round(float(it.mean()), 2), "poisson_iters_per_step_max": int(it.max()),
            "max_poisson_residual": max(r.poisson_residual for r in reps),
            "max_divergence": max(r.divergence_max for r in reps),
            "dt_last": last.dt, "t_end": last.t, "kinetic_energy_end": last.kinetic_energy,
            "inflow": last.inflow, "outflow": last.outflow,
            "p_finite": bool(np.isfinite(p[fl]).all())}


def octree_c3(ctx, with_cpu=True, flow_steps=200):
    """The cube-wake octree (4,294,942 leaves, levels 5..8, the cube's leaves
    InsideSolid): device setup (sort, validation, face table, operator), the
    SpMV roofline, a Jacobi-PCG solve to 1e-9 of a seeded mean-free
    right-hand side (OctreePoisson::solve) and the multigrid-preconditioned CG
    on the same system (hierarchy built beforehand, reported separately).  CPU side: the reference's
    faceNeighbors sweep over all leaves and one applyA, on one core."""
    t0 = time.perf_counter()
    t = nbl.build_tree(cases.C3_ORIGIN, cases.C3_EXTENT, 8, 5, cases.c3_predicate)
    k, l, c = t.leaves()
    host_build_s = time.perf_counter() - t0
    m = cases.c3_solid(c)
    d = nbl.Octree(ctx, k, l, m, cases.C3_ORIGIN, cases.C3_EXTENT, 8)
    setup_ms, _ = d.last_timing()
    n, nnz = d.counts()
    ms, by = d.profile(0, 50)
    tms, tby = d.profile(1, 50)
    sms, sby = d.profile(2, 20)
    zs = np.zeros(k.size)
    ones = np.ones(k.size)
    t1 = time.perf_counter()
    d.sensors(ones, zs, zs, zs, 1e-3, zs, zs)
    sens_call_s = time.perf_counter() - t1
    fl = m == 0
    b = cases.normal(11, k.size)
    b[~fl] = 0.0
    b[fl] -= b[fl].mean()
    xj, rep = d.solve(b, 1e-9, 20000)
    _, solve_ms = d.last_timing()
    levels, mg_build_ms = d.mg_levels()
    xm, mrep = d.mg_solve(b, 1e-9, 500)
    _, mg_ms = d.last_timing()
    variants = {}
    for name, prec in (("fp32_vcycle", nbl.PREC_MIXED), ("bf16_scratch", nbl.PREC_MIXED_BF16)):
        d.set_precision(prec)
        d.mg_solve(b, 1e-9, 500)  # first call of a form makes its FP32 operator copies
        xq, q = d.mg_solve(b, 1e-9, 500)
        _, qms = d.last_timing()
        res = float(np.linalg.norm((d.apply(xq) - b)[fl]) / np.linalg.norm(b))
        variants[name] = {"iterations": q.iterations, "residual": q.residual, "true_residual": res,
                          "converged": bool(q.converged), "solve_ms": round(qms, 3),
                          "speedup_vs_fp64_vcycle": round(mg_ms / max(qms, 1e-9), 3),
                          "rel_l2_vs_fp64_vcycle": rel(xq, xm)}
    flow = octree_flow(d, m, flow_steps) if flow_steps > 0 else None
    ncls = d.coef_classes()
    d.close()
    ach = by / (ms * 1e-3) / 1e9
    out = {"leaves": n, "couplings": nnz, "host_build_s": round(host_build_s, 3),
           "device_setup_ms": round(setup_ms, 3),
           "jacobi_pcg": {"tol": 1e-9, "iterations": rep.iterations, "residual": rep.residual,
                          "converged": bool(rep.converged), "solve_ms": round(solve_ms, 3),
                          "ms_per_iteration": round(solve_ms / max(rep.iterations, 1), 5)},
           "mg_pcg": {"tol": 1e-9, "levels": levels, "hierarchy_build_ms": round(mg_build_ms, 3),
                      "iterations": mrep.iterations, "residual": mrep.residual, "converged": bool(mrep.converged),
                      "solve_ms": round(mg_ms, 3), "speedup_vs_jacobi_pcg": round(solve_ms / max(mg_ms, 1e-9), 2),
                      "rel_l2_vs_jacobi_pcg": rel(xm, xj),
                      "mixed": variants,
                      "mixed_note": "the same flexible CG (FP64 residuals, dots and stopping test) with the "
                                    "V-cycle's levels in FP32, and with level 0's pre-smoothed iterate stored "
                                    "as BF16"},
           "spmv_roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(ach / HBM_PEAK_GBS, 4), "avg_launch_us": round(ms * 1e3, 3),
                             "bytes_per_launch": by, "coef_classes": ncls,
                             "note": "k_o_apply_dot; bytes: per leaf 8 B of coefficient classes and count, mask, "
                                     "x and y, per coupling a column index (and the FP64 coefficient of "
                                     "couplings beyond a row's sixth)" if ncls else
                                     "k_o_apply_dot; bytes: per leaf count, mask, x and y, per coupling a column "
                                     "index and an FP64 coefficient"},
           "transport_step": {"kernel": "k_oct_transport (octreeTransportStep, first-order upwind)",
                              "avg_launch_us": round(tms * 1e3, 3), "mleaves_per_s": round(n / tms / 1e3, 1),
                              "achieved_gbs": round(tby / tms / 1e6, 1), "bytes_per_launch": tby},
           "sensors": {"kernel": "k_oct_sensors (PhysicalSensorEvaluator::evaluate: gradients of u, v, w, p across "
                                 "level jumps, vorticity, strain, Q, wall shear, y+, truncation, stagnation)",
                       "avg_launch_us": round(sms * 1e3, 3), "mleaves_per_s": round(n / sms / 1e3, 1),
                       "achieved_gbs": round(sby / sms / 1e6, 1), "bytes_per_launch": sby,
                       "host_call_s": round(sens_call_s, 4),
                       "host_call_note": "nbl_octree_sensors with host arrays: PCIe copies of 6 inputs and 12 "
                                         "outputs included"}}
    if flow is not None:
        out["flow_steps"] = flow
        log(f"octree flow: {flow['steps']} steps, {flow['ms_per_step']} ms/step "
            f"({flow['mleaf_steps_per_s']} Mleaf-steps/s), projection {flow['projection_share']:.0%}, "
            f"{flow['poisson_iters_per_step_mean']} CG iterations/step, max div {flow['max_divergence']:.2e}")
    log(f"octree: {n} leaves, setup {setup_ms:.1f} ms, Jacobi-PCG {rep.iterations} iterations in {solve_ms:.1f} ms, "
        f"MG-PCG {mrep.iterations} iterations in {mg_ms:.1f} ms; FP32 cycle "
        f"{variants['fp32_vcycle']['iterations']} in {variants['fp32_vcycle']['solve_ms']:.1f} ms, BF16 "
        f"{variants['bf16_scratch']['iterations']} in {variants['bf16_scratch']['solve_ms']:.1f} ms")
    if with_cpu:
        import oracle_py
        kind = "reference" if oracle_py.available("ref") else "port"
        lib = oracle_py.OracleLib("ref" if kind == "reference" else "orc")
        core = OneCore().__enter__()
        rt = lib.octree(cases.C3_ORIGIN, cases.C3_EXTENT, 8)
        rt.refine_uniform(5)
        while True:
            rk, rl, _m = rt.dump()
            pred = cases.c3_predicate(rt.centers()) & (rl < 8)
            if not pred.any():
                break
            rt.refine_flagged_pass(rk[pred], rl[pred])
        rk, _rl, _m = rt.dump()
        for i in np.nonzero(cases.c3_solid(rt.centers()))[0]:
            rt.set_mask(int(i), 2)
        t1 = time.perf_counter()
        rt.face_table()
        conn_s = time.perf_counter() - t1
        xr = cases.normal(3, rt.n)
        t1 = time.perf_counter()
        rt.apply(xr)
        apply_s = time.perf_counter() - t1
        z = np.zeros(rt.n)
        t1 = time.perf_counter()
        rt.transport(xr, xr, z, xr, 1e-4)
        transport_s = time.perf_counter() - t1
        t1 = time.perf_counter()
        rt.sensors(xr, xr, z, xr, 1e-3, z, z)
        sensors_s = time.perf_counter() - t1
        rt.close()
        core.__exit__(None, None, None)
        out["cpu_baseline"] = {"kind": kind, "cores": 1, "connectivity_s": round(conn_s, 3),
                               "applyA_s": round(apply_s, 3), "transport_step_s": round(transport_s, 3),
                               "sensors_s": round(sensors_s, 3),
                               "est_solve_s": round(apply_s * rep.iterations, 1),
                               "sample": "faceNeighbors over every leaf, one OctreePoisson::applyA, one "
                                         "octreeTransportStep and one PhysicalSensorEvaluator::evaluate on the "
                                         "same tree; est_solve_s = applyA_s x the "
                                         "device Jacobi-PCG iteration count"}
        out["cpu_baseline"].update(core.record())
    return out


def batched(ctx, with_cpu=True, seed=4242, nuni=48, noct=8, concurrency=4):
    """48 independent uniform Poisson problems (n^3, n drawn from 32/64/96/128,
    all walls, i.i.d. N(0,1) right-hand sides projected to zero mean, MG-PCG to
    1e-8) and 8 octree problems (Jacobi-PCG and the multigrid CG to 1e-8): one
    solve after another versus nbl_*_solve_batch with `concurrency` solves in
    flight."""
    sizes, grids, rhs = cases.batch_uniform(nuni, seed)
    t0 = time.perf_counter()
    solvers = [nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=200) for g in grids]
    setup_s = time.perf_counter() - t0
    for p, b in zip(solvers, rhs):  # warm-up (first touches, code objects)
        p.solve(b)
    t0 = time.perf_counter()
    seq, seq_dev_ms = [], 0.0
    for p, b in zip(solvers, rhs):
        seq.append(p.solve(b))
        seq_dev_ms += p.last_timing()[0]
    seq_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    xs, reps = nbl.solve_batch(solvers, rhs, concurrency=concurrency)
    bat_s = time.perf_counter() - t0
    same = all(np.array_equal(x, y[0]) for x, y in zip(xs, seq))
    for p in solvers:
        p.close()
    # one batch per size class: every level kernel of a class is one launch
    # for all its problems (nbl_poisson_batch_*)
    classes = {n: [q for q, s in enumerate(sizes) if s == n] for n in (32, 64, 96, 128)}
    batches = {n: nbl.PoissonBatch(ctx, grids[idx[0]], len(idx), tol=1e-8, max_iters=200)
               for n, idx in classes.items() if idx}
    for n, pb in batches.items():  # warm-up
        pb.solve([rhs[q] for q in classes[n]])
    t0 = time.perf_counter()
    cls_dev_ms, cls_dev, cls_iters, cls_x = 0.0, {}, {}, {}
    for n, pb in batches.items():
        xs_c, reps_c = pb.solve([rhs[q] for q in classes[n]])
        ms, _v = pb.last_timing()
        cls_dev_ms += ms
        cls_dev[str(n)] = round(ms, 3)
        cls_iters[str(n)] = sorted({r.iterations for r in reps_c})
        for q, x in zip(classes[n], xs_c):
            cls_x[q] = (x, reps_c[classes[n].index(q)])
    cls_s = time.perf_counter() - t0
    # the class batches side by side, one host thread (and stream) each
    import threading
    cc_out = {}

    def run_class(n):
        cc_out[n] = batches[n].solve([rhs[q] for q in classes[n]])

    t0 = time.perf_counter()
    ths = [threading.Thread(target=run_class, args=(n,)) for n in batches]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    cc_s = time.perf_counter() - t0
    cc_same = all(np.array_equal(x, cls_x[q][0]) for n, (xs_c, _r) in cc_out.items()
                  for q, x in zip(classes[n], xs_c))
    for pb in batches.values():
        pb.close()
    cls_conv = all(r.converged for _, r in cls_x.values())
    cls_rel = max(rel(cls_x[q][0], seq[q][0]) for q in cls_x)
    cells = int(sum(g.cells for g in grids))
    out = {"uniform": {"problems": nuni, "cells": cells, "sizes": {str(n): sizes.count(n) for n in (32, 64, 96, 128)},
                       "setup_s": round(setup_s, 3), "sequential_s": round(seq_s, 4), "batched_s": round(bat_s, 4),
                       "speedup": round(seq_s / bat_s, 3), "concurrency": concurrency,
                       "all_converged": all(r.converged for r in reps),
                       "max_residual": max(r.residual for r in reps),
                       "iterations_mean": round(float(np.mean([r.iterations for r in reps])), 2),
                       "batch_equals_sequential": bool(same),
                       "mcells_solved_per_s": round(cells / bat_s / 1e6, 2),
                       "sequential_device_ms": round(seq_dev_ms, 3),
                       "size_class_batches": {
                           "note": "one nbl_poisson_batch per size class (14/12/12/10 problems): one launch per "
                                   "level kernel for all problems of a class, per-problem scalars and stopping",
                           "host_arrays_s": round(cls_s, 4), "speedup_vs_sequential": round(seq_s / cls_s, 3),
                           "speedup_vs_concurrent": round(bat_s / cls_s, 3),
                           "device_ms": round(cls_dev_ms, 3), "device_ms_per_class": cls_dev,
                           "device_speedup_vs_sequential": round(seq_dev_ms / max(cls_dev_ms, 1e-9), 3),
                           "classes_concurrent_s": round(cc_s, 4),
                           "classes_concurrent_speedup_vs_sequential": round(seq_s / cc_s, 3),
                           "classes_concurrent_equal": bool(cc_same),
                           "iterations_per_class": cls_iters, "all_converged": bool(cls_conv),
                           "max_rel_l2_vs_sequential": cls_rel}}}
    trees = cases.batch_octrees(noct, seed + 99)
    octs, orhs = [], []
    for q, (k, lv) in enumerate(trees):
        octs.append(nbl.Octree(ctx, k, lv, np.zeros(k.size, np.uint8), (0, 0, 0), (1, 1, 1), 8))
        b = cases.normal(seed + 500 + q, k.size)
        orhs.append(b - b.mean())
    t0 = time.perf_counter()
    oseq = [o.solve(b, 1e-8, 20000) for o, b in zip(octs, orhs)]
    oseq_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    oxs, oreps = nbl.octree_solve_batch(octs, orhs, tol=1e-8, max_iters=20000, concurrency=concurrency)
    obat_s = time.perf_counter() - t0
    osame = all(np.array_equal(x, y[0]) for x, y in zip(oxs, oseq))
    # multigrid CG on the same problems; hierarchies built before the timing
    for o in octs:
        o.mg_levels()
    t0 = time.perf_counter()
    mseq = [o.mg_solve(b, 1e-8, 500) for o, b in zip(octs, orhs)]
    mseq_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    mxs, mreps = nbl.octree_solve_batch(octs, orhs, tol=1e-8, max_iters=500, concurrency=concurrency, mg=True)
    mbat_s = time.perf_counter() - t0
    msame = all(np.array_equal(x, y[0]) for x, y in zip(mxs, mseq))
    for o in octs:
        o.close()
    out["octree"] = {"problems": noct, "leaves": [int(k.size) for k, _ in trees],
                     "sequential_s": round(oseq_s, 4), "batched_s": round(obat_s, 4),
                     "speedup": round(oseq_s / obat_s, 3), "all_converged": all(r.converged for r in oreps),
                     "iterations": [r.iterations for r in oreps], "batch_equals_sequential": bool(osame),
                     "mg_pcg": {"sequential_s": round(mseq_s, 4), "batched_s": round(mbat_s, 4),
                                "speedup": round(mseq_s / mbat_s, 3),
                                "all_converged": all(r.converged for r in mreps),
                                "iterations": [r.iterations for r in mreps],
                                "batch_equals_sequential": bool(msame),
                                "speedup_vs_jacobi_pcg_batched": round(obat_s / mbat_s, 2)}}
    if with_cpu:
        import oracle_py
        kind = "reference" if oracle_py.available("ref") else "port"
        lib = oracle_py.OracleLib("ref" if kind == "reference" else "orc")
        sample = [q for q, n in enumerate(sizes) if n in (32, 64)][:6]
        with OneCore() as core:
            t0 = time.perf_counter()
            cpu_cells = 0
            for q in sample:
                lib.mg(grids[q], rhs[q], 1e-8, 200)
                cpu_cells += grids[q].cells
            el = time.perf_counter() - t0
        out["cpu_baseline"] = {"kind": kind, "cores": 1, "sample_problems": len(sample),
                               "sample_cells": cpu_cells, "seconds": round(el, 3),
                               "mcells_solved_per_s": round(cpu_cells / el / 1e6, 4),
                               "sample": "MgPoisson::solve to 1e-8 on the batch's 32^3 and 64^3 problems (up to 6)"}
        out["cpu_baseline"].update(core.record())
    return out


def served(ctx, window_s, rate, nruns=32, slots=4, seed=777, sustain_window_s=10.0, precision=None):
    """Poisson arrivals at `rate` requests/s over `window_s` seconds, each one
    RK3 step of a uniformly drawn run; latency = submit to completion.  The
    same server then takes a second stream at 70% of the request rate it
    completed in the first (a load it sustains), over `sustain_window_s`."""
    import threading
    runs = cases.served_runs(nruns, seed)
    srv = nbl.Server(ctx, slots=slots, hbm_budget=200 << 30)
    t0 = time.perf_counter()
    opts = nbl.solver_opts(precision=precision) if precision is not None else None
    ids = [cases.admit_served(srv, r, opts) for r in runs]
    cells = [r["cells"] for r in runs]
    admit_s = time.perf_counter() - t0
    used, _ = srv.footprint()

    def stream(rate_, window_, useed):
        u = cases.uniform01(useed, 2 * int(rate_ * window_ * 2) + 16)
        arrivals, pick, t = [], [], 0.0
        q = 0
        while True:
            t += -np.log(max(float(u[2 * q]), 1e-300)) / rate_
            if t > window_:
                break
            arrivals.append(t)
            pick.append(int(float(u[2 * q + 1]) * nruns) % nruns)
            q += 1
        tickets = []
        lat, done_cells = [], []
        lock = threading.Lock()
        feeding_done = [False]

        def collector():
            k = 0
            while True:
                with lock:
                    if k >= len(tickets) and feeding_done[0]:
                        return
                    have = k < len(tickets)
                    if have:
                        tk, rid = tickets[k]
                if not have:
                    time.sleep(0.0005)
                    continue
                _rep, ms = srv.wait(tk)
                lat.append(ms)
                done_cells.append(cells[rid])
                k += 1

        th = threading.Thread(target=collector)
        th.start()
        start = time.perf_counter()
        for at, rid in zip(arrivals, pick):
            dly = start + at - time.perf_counter()
            if dly > 0:
                time.sleep(dly)
            tk = srv.submit(ids[rid])
            with lock:
                tickets.append((tk, rid))
        with lock:
            feeding_done[0] = True
        th.join()
        elapsed = time.perf_counter() - start
        lat = np.array(lat)
        return {"requests": len(lat), "window_s": window_, "rate_per_s": round(rate_, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3), "drain_s": round(elapsed - window_, 3),
                "completed_per_s": round(len(lat) / elapsed, 2),
                "mcell_steps_per_s": round(float(np.sum(done_cells)) / elapsed / 1e6, 3)}

    out = stream(rate, window_s, seed + 1)
    out.update({"runs": nruns, "slots": slots, "cells_min": int(min(cells)), "cells_max": int(max(cells)),
                "cells_mean": int(np.mean(cells)), "admit_s": round(admit_s, 3), "hbm_bytes": int(used),
                "octree_runs": sum(r["kind"] == "octree" for r in runs),
                "families": "Taylor-Green n^3 and cube channel 12r x 4r x 4r (uniform MAC), cube-wake octrees "
                            "(finest level 7 or 8, collocated octree step); Re U[500,3000]"})
    if sustain_window_s > 0:
        out["sustained"] = stream(0.7 * out["completed_per_s"], sustain_window_s, seed + 2)
        out["sustained"]["note"] = "70% of the request rate the server completed under the overload above"
    srv.close()
    return out


def tgv_c1(ctx, a, steps=100):
    """The 64^3 periodic Taylor-Green vortex (Re 1600, CFL 0.5, WENO-5, singular
    MG-PCG to 1e-8), 100 steps from the analytic state in FP64 and in the mixed
    form, with one step of the same case on one host core."""
    g, cfg = cases.taylor_green(64)
    st = cases.taylor_green_state(g)
    res = {"workload": "Taylor-Green 64^3 periodic, Re 1600, CFL 0.5, WENO-5, MG-PCG tol 1e-8 (singular)",
           "cells": g.cells, "steps": steps}
    state = None
    for name, prec in (("fp64", nbl.PREC_FP64), ("mixed", nbl.PREC_MIXED)):
        reps, ms, it, el, stt = run_steps(ctx, g, cfg, st, prec, 3, steps)
        res[name] = {"value": round(g.cells * steps / el / 1e6, 3), "ms_per_step": round(el / steps * 1e3, 4),
                     "mg_pcg_solve_ms_median": round(statistics.median(ms), 4),
                     "poisson_iters_per_solve_mean": round(float(np.mean(it)), 2),
                     "max_poisson_residual": max(r["poisson_residual"] for r in reps)}
        if state is None:
            state = stt
        else:
            res[name]["rel_l2_vs_fp64"] = {k: rel(x, y) for k, x, y in zip("uvwp", stt[:4], state[:4])}
        log(f"tgv64 {name}: {res[name]['value']} Mcell-steps/s")
    res["roofline"] = roofline(ctx, g, reps=200, kernel=0)
    res["roofline"]["note"] = ("64^3 singular: the finest-level sweep of the one-kernel-per-sweep V-cycle (6.8 MB "
                               "per launch at FP64), latency-bound and Infinity-Cache resident")
    res["roofline_mixed"] = roofline(ctx, g, reps=200, kernel=0, precision=nbl.PREC_MIXED)
    if not a.no_cpu and a.cpu_steps > 0:
        res["cpu_baseline"] = cpu_baseline(g, cfg, st, 8)
    return res


def host_info():
    """The GPU host's CPU as the CPU baselines record it: model, logical CPUs
    (nproc) and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = []
    return {"cpu_model": model, "nproc": os.cpu_count(), "allowed_cpus": len(allowed)}


class OneCore:
    """Pin this process to one CPU (the lowest it may use) for a single-thread
    timing, as `taskset -c <cpu>` would; the record names the CPU."""

    def __enter__(self):
        self.cpu = None
        try:
            self.old = os.sched_getaffinity(0)
            self.cpu = min(self.old)
            os.sched_setaffinity(0, {self.cpu})
        except (AttributeError, OSError):
            self.old = None
        return self

    def __exit__(self, *exc):
        if self.old is not None:
            os.sched_setaffinity(0, self.old)
        return False

    def record(self):
        out = host_info()
        out["affinity"] = [self.cpu] if self.cpu is not None else "unpinned"
        return out


def cpu_baseline(g, cfg, st, steps):
    import oracle_py
    kind = "reference" if oracle_py.available("ref") else "port"
    lib = oracle_py.OracleLib("ref" if kind == "reference" else "orc")
    with OneCore() as core:
        s = lib.ns(g, cfg)
        s.set_state(*st)
        t0 = time.perf_counter()
        for q in range(steps):
            s.step()
            log(f"cpu baseline step {q + 1}/{steps}")
        el = time.perf_counter() - t0
        s.close()
    out = {"value": round(g.cells * steps / el / 1e6, 5), "unit": "Mcell-steps/s", "cores": 1,
           "kind": kind, "seconds_per_step": round(el / steps, 3),
           "sample": f"{steps} steps of the same workload from the same initial state, 1 thread"}
    out.update(core.record())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-octree", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--octree-flow-steps", type=int, default=200,
                    help="flow steps on the cube-wake octree (0: skip)")
    ap.add_argument("--octree-flow-only", action="store_true",
                    help="only the cube-wake octree's flow steps (for profiler runs)")
    ap.add_argument("--no-tgv", action="store_true")
    ap.add_argument("--served-window", type=float, default=60.0,
                    help="seconds of Poisson arrivals in the served-stream measurement (0: skip)")
    ap.add_argument("--served-rate", type=float, default=200.0)
    ap.add_argument("--profile-only", action="store_true",
                    help="only the timed FP64 steps (for profiler runs)")
    ap.add_argument("--workload", choices=("cube192", "tgv64"), default="cube192",
                    help="with --profile-only: the cube channel or the 64^3 Taylor-Green vortex")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline kernel loops and a 256 MiB copy (for profiler and counter runs)")
    ap.add_argument("--roofline-singular", action="store_true",
                    help="only the FP64 sweep on the singular (all-wall) 256^3 level and a 256 MiB copy")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(independent runs shard across GPUs without exchange: one server per GPU)", file=sys.stderr)
        sys.exit(2)

    g, cfg = cases.cube_channel()
    st = cases.cube_channel_state(g)
    ctx = nbl.Ctx(0)
    if a.profile_only and a.workload == "tgv64":
        tg, tcfg = cases.taylor_green(64)
        _r, _ms, _it, el, _s = run_steps(ctx, tg, tcfg, cases.taylor_green_state(tg), nbl.PREC_FP64, a.warmup, a.steps)
        ctx.close()
        print(json.dumps({"workload": "tgv64", "steps": a.steps, "ms_per_step": round(el / a.steps * 1e3, 4),
                          "value": round(tg.cells * a.steps / el / 1e6, 3)}), flush=True)
        return
    if a.octree_flow_only:
        t = nbl.build_tree(cases.C3_ORIGIN, cases.C3_EXTENT, 8, 5, cases.c3_predicate)
        k, l, c = t.leaves()
        m = cases.c3_solid(c)
        d = nbl.Octree(ctx, k, l, m, cases.C3_ORIGIN, cases.C3_EXTENT, 8)
        fo = octree_flow(d, m, a.octree_flow_steps, warmup=a.warmup)
        d.close()
        ctx.close()
        print(json.dumps({"octree_flow": fo}), flush=True)
        return
    if a.roofline_singular:
        out = {"singular_fp64_sweep": roofline(ctx, grid256(singular=True), key="256x256x256"),
               "copy_256MiB_ms": ctx.probe_copy(256 << 20, 20)}
        ctx.close()
        print(json.dumps(out), flush=True)
        return
    if a.roofline_only:
        # the kernels of the roofline figures, each its own template instance
        # (so a profiler's per-kernel summary of this run is theirs alone)
        big = grid256()
        tg = cases.taylor_green(64)[0]
        out = {"roofline": roofline(ctx, big, key="256x256x256_outlet"),
               "fp32_sweep": roofline(ctx, big, precision=nbl.PREC_MIXED, key="256x256x256_outlet_fp32"),
               "spmv_fp64": roofline(ctx, big, kernel=1, key="256x256x256_outlet_spmv"),
               "tgv64_down_fp64": roofline(ctx, tg, kernel=4, key="tgv64_down_fp64"),
               "tgv64_sweep_fp64": roofline(ctx, tg, kernel=0),
               "copy_256MiB_ms": ctx.probe_copy(256 << 20, 20)}
        ctx.close()
        print(json.dumps(out), flush=True)
        return
    log(f"workload 192x64x64 cube channel, {g.cells} cells, warmup {a.warmup}, steps {a.steps}")
    sweep = {}
    reps, solve_ms, iters, el, state = run_steps(ctx, g, cfg, st, nbl.PREC_FP64, a.warmup, a.steps, timing=sweep)
    ms_step = el / a.steps * 1e3
    value = g.cells * a.steps / el / 1e6
    log(f"fp64: {ms_step:.3f} ms/step, {value:.2f} Mcell-steps/s")
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded splitmix64 perturbation of the inflow)",
        "config": {"workload": "cube-channel 192x64x64, Re_h 500, WENO-5, SSP-RK3, MG-PCG tol 1e-8",
                   "cells": g.cells, "precision": "fp64"},
        "mg_pcg_solve_ms_median": round(statistics.median(solve_ms), 4),
        "poisson_iters_per_solve_mean": round(float(np.mean(iters)), 2),
        "max_poisson_residual": max(r["poisson_residual"] for r in reps),
        "served_p50_ms": None, "served_p99_ms": None,
    }
    if a.profile_only:
        out["roofline_cube192_infinity_cache"] = roofline_insitu(sweep, g)
        print(json.dumps(out), flush=True)
        return
    if not a.no_mixed:
        mreps, mms, mit, mel, mstate = run_steps(ctx, g, cfg, st, nbl.PREC_MIXED, a.warmup, a.steps)
        mval = g.cells * a.steps / mel / 1e6
        cd = np.array([r["cd"] for r in reps])
        mcd = np.array([r["cd"] for r in mreps])
        cl = np.array([r["cl"] for r in reps])
        mcl = np.array([r["cl"] for r in mreps])
        out["mixed_fp32_vcycle"] = {
            "value": round(mval, 3), "ms_per_step": round(mel / a.steps * 1e3, 4),
            "speedup_vs_fp64": round(mval / value, 3),
            "mg_pcg_solve_ms_median": round(statistics.median(mms), 4),
            "poisson_iters_per_solve_mean": round(float(np.mean(mit)), 2),
            "max_poisson_residual": max(r["poisson_residual"] for r in mreps),
            "rel_l2_vs_fp64": {k: rel(x, y) for k, x, y in zip("uvwp", mstate[:4], state[:4])},
            "cd_max_rel_err": float(np.max(np.abs(mcd - cd) / np.maximum(np.abs(cd), 1e-30))),
            "cl_max_abs_err": float(np.max(np.abs(mcl - cl))),
        }
        log(f"mixed: {mval:.2f} Mcell-steps/s")
    # The reference's loop structure on the GPU, one kernel per loop with its
    # CgPoisson(1e-6, 500) coarsest solve: the unfused FP64 baseline
    lreps, lms, lit, lel, lstate = run_steps(ctx, g, cfg, st, nbl.PREC_FP64, a.warmup, a.steps, loops=True)
    lval = g.cells * a.steps / lel / 1e6
    out["fp64_reference_loops"] = {
        "value": round(lval, 3), "ms_per_step": round(lel / a.steps * 1e3, 4),
        "mg_pcg_solve_ms_median": round(statistics.median(lms), 4),
        "poisson_iters_per_solve_mean": round(float(np.mean(lit)), 2),
        "max_poisson_residual": max(r["poisson_residual"] for r in lreps),
        "rel_l2_vs_fp64": {k: rel(x, y) for k, x, y in zip("uvwp", lstate[:4], state[:4])},
        "fused_fp64_speedup": round(value / lval, 3),
        "note": "one kernel per loop of multigrid.cpp / poisson.cpp (smooth, residual, restrict, prolong, "
                "removeMean, every dot and update), CgPoisson(1e-6, 500) on the coarsest level"}
    if "mixed_fp32_vcycle" in out:
        out["mixed_fp32_vcycle"]["speedup_vs_fp64_reference_loops"] = round(out["mixed_fp32_vcycle"]["value"] / lval, 3)
    log(f"fp64 reference loops: {lval:.2f} Mcell-steps/s")
    if not a.no_tgv:
        out["tgv64_c1"] = tgv_c1(ctx, a)
    big = grid256()
    out["roofline"] = roofline(ctx, big, key="256x256x256_outlet")
    out["roofline"]["workload_note"] = ("the channel's dominant kernel at HBM scale: 256^3 walls + outlet "
                                        "(non-singular, as the channel's system), 436 MB per launch")
    out["roofline_256cubed"] = {
        "fp32_sweep": roofline(ctx, big, precision=nbl.PREC_MIXED, key="256x256x256_outlet_fp32"),
        "spmv_fp64": roofline(ctx, big, kernel=1, key="256x256x256_outlet_spmv"),
        "singular_fp64_sweep": roofline(ctx, grid256(singular=True), key="256x256x256"),
    }
    out["roofline_cube192_infinity_cache"] = {
        "in_step_stamps": roofline_insitu(sweep, g),
        "isolated_events": roofline(ctx, g, reps=200, key="cube192x64x64"),
        "note": "192x64x64: the sweep's 20 MB and the solve's ~170 MB working set are resident in the 256 MB "
                "Infinity Cache, so these are not HBM figures; the stamps leave out every launch's dispatch cost"}
    out["coarse_solve"] = coarse_solve(ctx, g)
    rb = out["roofline_256cubed"]
    log(f"256^3: fp64 sweep {out['roofline']['frac']}, fp32 sweep {rb['fp32_sweep']['frac']}, "
        f"spmv {rb['spmv_fp64']['frac']}, singular fp64 sweep {rb['singular_fp64_sweep']['frac']}")
    if not a.no_octree:
        out["octree_c3"] = octree_c3(ctx, with_cpu=not a.no_cpu, flow_steps=a.octree_flow_steps)
    if not a.no_batch:
        out["batched_solves"] = batched(ctx, with_cpu=not a.no_cpu)
        log(f"batched: {out['batched_solves']['uniform']['speedup']}x uniform, "
            f"{out['batched_solves']['octree']['speedup']}x octree")
    if a.served_window > 0:
        # the served form: FP32 V-cycle inside the FP64 flexible CG
        sv = served(ctx, a.served_window, a.served_rate, precision=nbl.PREC_MIXED)
        sv["precision"] = "mixed (FP32 V-cycle, FP64 outer CG)"
        sv["speedup_vs_fp64_reference_loops"] = round(sv["mcell_steps_per_s"] / lval, 3)
        sv["speedup_vs_fused_fp64"] = round(sv["mcell_steps_per_s"] / value, 3)
        out["served"] = sv
        out["served_p50_ms"], out["served_p99_ms"] = sv["p50_ms"], sv["p99_ms"]
        log(f"served: {sv['requests']} requests, p50 {sv['p50_ms']} ms, p99 {sv['p99_ms']} ms")
        if "sustained" in sv:
            ss = sv["sustained"]
            out["served_sustained_p50_ms"], out["served_sustained_p99_ms"] = ss["p50_ms"], ss["p99_ms"]
            log(f"served at {ss['rate_per_s']} req/s: p50 {ss['p50_ms']} ms, p99 {ss['p99_ms']} ms")
    ctx.close()
    if not a.no_cpu and a.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(g, cfg, st, a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
