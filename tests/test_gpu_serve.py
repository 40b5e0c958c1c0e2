"""Serving loop and batched solves through the C-ABI.

* a step served while other runs' steps are in flight returns the same report
  and state, bit for bit, as the same step run alone (the kernels are
  deterministic and a run's arithmetic never depends on other streams);
* requests for a busy run are served in order and later requests for idle
  runs overtake them;
* octree flow runs are served beside uniform ones, with the same guarantee;
* admission refuses a run whose device footprint exceeds the budget;
* a batch of solves equals the same solves one by one, bit for bit.
"""
import threading

import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu


def small_runs():
    out = []
    g, cfg = cases.taylor_green(16)
    out.append((g, cfg, cases.taylor_green_state(g)))
    g, cfg = cases.cube_channel(nx=48, ny=16, nz=16, res=4)
    out.append((g, cfg, cases.cube_channel_state(g)))
    g, cfg = cases.taylor_green(24)
    out.append((g, cfg, cases.taylor_green_state(g)))
    g, cfg = cases.cube_channel(nx=32, ny=16, nz=16, res=4)
    out.append((g, cfg, cases.cube_channel_state(g, seed=5)))
    return out


def test_served_steps_equal_alone(ctx):
    runs = small_runs()
    steps = 3
    alone = []
    for g, cfg, st in runs:
        r = nbl.Run(ctx, g, cfg)
        r.set_state(*st)
        reps = [r.step().as_dict() for _ in range(steps)]
        alone.append((reps, r.get_state()))
        r.close()
    srv = nbl.Server(ctx, slots=4)
    try:
        ids = []
        for g, cfg, st in runs:
            rid = srv.admit(g, cfg)
            srv.run(rid, g).set_state(*st)
            ids.append(rid)
        # all requests at once, interleaved across runs
        tickets = [(rid, srv.submit(rid)) for _ in range(steps) for rid in ids]
        got = {rid: [] for rid in ids}
        for rid, t in tickets:
            rep, ms = srv.wait(t)
            assert ms > 0.0
            got[rid].append(rep.as_dict())
        for (g, _cfg, _st), rid, (reps, state) in zip(runs, ids, alone):
            assert got[rid] == reps
            s2 = srv.run(rid, g).get_state()
            for a, b in zip(s2[:4], state[:4]):
                assert np.array_equal(a, b)
    finally:
        srv.close()


def test_served_octree_runs_equal_alone(ctx):
    """Octree flow runs (cube-wake trees of finest level 6 and 7) served beside
    a uniform run: every served step report and the final state equal the
    run stepped alone; handles of the wrong kind are refused."""
    runs = [cases.cube_wake_run(6, 3.0, 800.0, 11), cases.cube_wake_run(7, 1.5, 1500.0, 12)]
    g, cfg = cases.taylor_green(16)
    runs.append(dict(name="tgv16", kind="uniform", cells=g.cells, grid=g, cfg=cfg,
                     state=cases.taylor_green_state(g)))
    steps = 3
    srv = nbl.Server(ctx, slots=3)
    try:
        ids = [cases.admit_served(srv, r) for r in runs]
        with pytest.raises(nbl.NblError):
            srv.run(ids[0], g)
        with pytest.raises(nbl.NblError):
            srv.octree_flow(ids[2])
        tickets = [(q, srv.submit(rid)) for _ in range(steps) for q, rid in enumerate(ids)]
        got = [[] for _ in runs]
        for q, t in tickets:
            got[q].append(srv.wait(t)[0])
        states = [cases.served_state(srv, rid, r) for rid, r in zip(ids, runs)]
    finally:
        srv.close()
    for q, r in enumerate(runs):
        reps, ref = cases.step_alone(ctx, r, steps)
        for a, b in zip(got[q], reps):
            if r["kind"] == "octree":
                assert (a.step, a.t, a.dt, a.divergence_max, a.poisson_iters, a.poisson_residual) == \
                       (b.step, b.t, b.dt, b.divergence_max, b.poisson_iters, b.poisson_residual)
                assert a.mass_error == b.inflow - b.outflow and a.cell_count == r["cells"]
                assert a.cfl == r["cfg"].cfl and a.accepted == 1
            else:
                assert a.as_dict() == b.as_dict()
        for fa, fb in zip(states[q], ref):
            assert np.array_equal(fa, fb), r["name"]


def test_server_admission_budget(ctx):
    g, cfg = cases.taylor_green(16)
    srv = nbl.Server(ctx, slots=1, hbm_budget=1 << 20)  # 1 MiB: too small for any run
    try:
        with pytest.raises(nbl.NblError) as e:
            srv.admit(g, cfg)
        assert e.value.code == -4  # NBL_ERR_CAPACITY
        assert srv.footprint()[0] == 0
    finally:
        srv.close()


def test_server_concurrent_submitters(ctx):
    """Several host threads submitting and waiting at once."""
    runs = small_runs()
    srv = nbl.Server(ctx, slots=3)
    try:
        ids = []
        for g, cfg, st in runs:
            rid = srv.admit(g, cfg)
            srv.run(rid, g).set_state(*st)
            ids.append(rid)
        errors = []

        def client(rid):
            try:
                for _ in range(2):
                    rep, _ms = srv.wait(srv.submit(rid))
                    assert rep.poisson_residual <= 1e-7
            except Exception as ex:  # surfaced below
                errors.append(ex)

        th = [threading.Thread(target=client, args=(rid,)) for rid in ids]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors
        with pytest.raises(nbl.NblError):
            srv.wait(123456789)
    finally:
        srv.close()


def test_batch_equals_sequential(ctx):
    grids = [nbl.Grid(n, n, n, 1.0 / n) for n in (16, 24, 32, 16)]
    rhs = []
    for q, g in enumerate(grids):
        b = cases.normal(100 + q, g.cells)
        rhs.append(b - b.mean())
    seq = []
    for g, b in zip(grids, rhs):
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=200)
        seq.append(ps.solve(b))
        ps.close()
    solvers = [nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=200) for g in grids]
    try:
        xs, reps = nbl.solve_batch(solvers, rhs, concurrency=3)
    finally:
        for p in solvers:
            p.close()
    for (x0, r0), x1, r1 in zip(seq, xs, reps):
        assert r1.converged and r1.iterations == r0.iterations
        assert np.array_equal(x0, x1)


@pytest.mark.parametrize("mg", [False, True], ids=["jacobi", "mg"])
def test_octree_batch_equals_sequential(ctx, orc, mg):
    """Batched octree solves (Jacobi-PCG or the multigrid CG, each octree on
    its own stream) return the same solutions bit for bit as one solve after
    another."""
    from octree_trees import adaptive_octree, uniform_octree
    trees = [uniform_octree(orc, 4), adaptive_octree(orc), uniform_octree(orc, 5)]
    octs, rhs = [], []
    for q, t in enumerate(trees):
        k, l, m = t.dump()
        octs.append(nbl.Octree(ctx, k, l, m, t.origin, t.extent, t.max_level))
        fl = m == 0
        b = cases.normal(300 + q, t.n)
        b[~fl] = 0.0
        b[fl] -= b[fl].mean()
        rhs.append(b)
    try:
        solve = (lambda o, b: o.mg_solve(b, 1e-9, 300)) if mg else (lambda o, b: o.solve(b, 1e-9, 5000))
        seq = [solve(o, b) for o, b in zip(octs, rhs)]
        xs, reps = nbl.octree_solve_batch(octs, rhs, tol=1e-9, max_iters=300 if mg else 5000,
                                          concurrency=2, mg=mg)
    finally:
        for o in octs:
            o.close()
    for (x0, r0), x1, r1 in zip(seq, xs, reps):
        assert r1.converged and r1.iterations == r0.iterations
        assert np.array_equal(x0, x1)
