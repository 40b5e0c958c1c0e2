"""Serving loop and batched solves through the C-ABI.

* a step served while other runs' steps are in flight returns the same report
  and state, bit for bit, as the same step run alone (the kernels are
  deterministic and a run's arithmetic never depends on other streams);
* requests for a busy run are served in order and later requests for idle
  runs overtake them;
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
