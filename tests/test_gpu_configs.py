"""The batched-solve and served-stream workloads at their full sizes, through
the C-ABI batch and server entry points.

* Batched solves: the 48 seeded n^3 problems (n in 32/64/96/128, all walls,
  N(0,1) right-hand sides projected to zero mean, tolerance 1e-8) and 8 octrees
  of 0.5-2M leaves.  Every problem's true relative residual, recomputed here
  from the returned solution with the oracle's operator (uniform) or the
  device operator that is bit-exact to it (octree), is <= 1e-8 and its fluid
  mean is zero; the batch returns bit for bit what one solve after another
  returns; the 32^3 and 64^3 problems agree with the oracle's MgPoisson.
* Served stream: 32 runs of 0.3-4M cells (Taylor-Green boxes and cube
  channels, Re 500-3000) admitted to one server, two interleaved step requests
  each; every run ends bit-identical to the same run stepped alone.
"""
import numpy as np
import pytest

import cases
import nbl

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_batched_uniform_solves_at_size(ctx, orc):
    sizes, grids, rhs = cases.batch_uniform(48, 4242)
    assert sorted(set(sizes)) == [32, 64, 96, 128]
    solvers = [nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=200) for g in grids]
    try:
        seq = [p.solve(b) for p, b in zip(solvers, rhs)]
        xs, reps = nbl.solve_batch(solvers, rhs)
    finally:
        for p in solvers:
            p.close()
    checked = 0
    for q, (g, b, x, r) in enumerate(zip(grids, rhs, xs, reps)):
        assert np.array_equal(x, seq[q][0]), q
        assert (r.iterations, r.residual, r.converged) == (seq[q][1].iterations, seq[q][1].residual,
                                                           seq[q][1].converged)
        assert r.converged
        res = np.linalg.norm(b - orc.apply(g, x)) / np.linalg.norm(b)
        assert res <= 1e-8 * (1 + 1e-6), (q, res)
        assert abs(x.mean()) <= 1e-12 * np.abs(x).max(), q
        if g.nx <= 64:
            xo, ro = orc.mg(g, b, 1e-8, 200)
            assert abs(r.iterations - ro.iterations) <= 1, q
            assert rel(x, xo) < 1e-7, q
            checked += 1
    print(f"48 problems, {sum(g.cells for g in grids)} cells; {checked} compared with the oracle", flush=True)
    assert checked >= 10


def test_batched_uniform_solves_in_size_classes(ctx, orc):
    """The same 48 problems as batches, one per size class (one launch per
    level kernel for all problems of a class): every problem converges to
    1e-8 (true residual recomputed with the oracle's operator), its fluid
    mean is zero, and it agrees with the same problem solved alone (iterations
    within one, solution within rounding of the tolerance)."""
    sizes, grids, rhs = cases.batch_uniform(48, 4242)
    for n in (32, 64, 96, 128):
        idx = [q for q, s in enumerate(sizes) if s == n]
        g = grids[idx[0]]
        with nbl.PoissonBatch(ctx, g, len(idx), tol=1e-8, max_iters=200) as pb:
            xs, reps = pb.solve([rhs[q] for q in idx])
        ps = nbl.Poisson(ctx, g, kind=0, tol=1e-8, max_iters=200)
        try:
            for x, r, q in zip(xs, reps, idx):
                b = rhs[q]
                assert r.converged, q
                res = np.linalg.norm(b - orc.apply(g, x)) / np.linalg.norm(b)
                assert res <= 1e-8 * (1 + 1e-6), (q, res)
                assert abs(x.mean()) <= 1e-12 * np.abs(x).max(), q
                x1, r1 = ps.solve(b)
                assert abs(r.iterations - r1.iterations) <= 1, q
                assert rel(x, x1) < 1e-7, q
        finally:
            ps.close()
        print(f"{n}^3: {len(idx)} problems in one batch, iterations {[r.iterations for r in reps]}", flush=True)


@pytest.mark.parametrize("mg", [False, True], ids=["jacobi_pcg", "mg_pcg"])
def test_batched_octree_solves_at_size(ctx, mg):
    trees = cases.batch_octrees(8, 4242 + 99)
    octs, rhs = [], []
    try:
        for q, (k, lv) in enumerate(trees):
            assert 500_000 <= k.size <= 2_100_000
            octs.append(nbl.Octree(ctx, k, lv, np.zeros(k.size, np.uint8), (0, 0, 0), (1, 1, 1), 8))
            b = cases.normal(4242 + 500 + q, k.size)
            rhs.append(b - b.mean())
        maxit = 500 if mg else 20000
        seq = [(o.mg_solve(b, 1e-8, maxit) if mg else o.solve(b, 1e-8, maxit)) for o, b in zip(octs, rhs)]
        xs, reps = nbl.octree_solve_batch(octs, rhs, tol=1e-8, max_iters=maxit, mg=mg)
        for q, (o, b, x, r) in enumerate(zip(octs, rhs, xs, reps)):
            assert np.array_equal(x, seq[q][0]), q
            assert r.iterations == seq[q][1].iterations and r.converged
            res = np.linalg.norm(b - o.apply(x)) / np.linalg.norm(b)
            assert res <= 1e-8 * (1 + 1e-6), (q, res)
            assert abs(x.mean()) <= 1e-12 * np.abs(x).max(), q
        print(f"8 octrees {[k.size for k, _ in trees]}: iterations {[r.iterations for r in reps]}", flush=True)
    finally:
        for o in octs:
            o.close()


def test_served_stream_at_size_equals_alone(ctx):
    """C5's 32 seeded runs (Taylor-Green boxes, cube channels and cube-wake
    octrees of 0.3-4M cells or leaves) admitted to one server, two requests
    each submitted together so that different runs' steps overlap on the
    device: every served step and final state equals the run stepped alone."""
    runs = cases.served_runs(32, 777)
    cells = [r["cells"] for r in runs]
    assert min(cells) >= 300_000 and max(cells) <= 4_000_000
    assert {r["kind"] for r in runs} == {"uniform", "octree"}
    opts = nbl.solver_opts(nbl.PREC_MIXED)
    srv = nbl.Server(ctx, slots=4, hbm_budget=200 << 30)
    try:
        ids = [cases.admit_served(srv, r, opts) for r in runs]
        tickets = [srv.submit(rid) for _ in range(2) for rid in ids]
        served = [srv.wait(t)[0] for t in tickets]
        states = [cases.served_state(srv, rid, r) for rid, r in zip(ids, runs)]
    finally:
        srv.close()
    for q, r in enumerate(runs):
        alone, ref = cases.step_alone(ctx, r, 2, opts)
        mine = [served[q], served[len(runs) + q]]
        for a, b in zip(mine, alone):
            assert (a.step, a.poisson_iters, a.dt, a.poisson_residual) == \
                   (b.step, b.poisson_iters, b.dt, b.poisson_residual), r["name"]
            if r["kind"] == "uniform":
                assert (a.cd, a.cl) == (b.cd, b.cl), r["name"]
            else:
                assert a.cell_count == r["cells"] and a.divergence_max == b.divergence_max, r["name"]
                assert a.mass_error == b.inflow - b.outflow, r["name"]
        for fa, fb in zip(states[q], ref):
            assert np.array_equal(fa, fb), r["name"]
    print(f"32 runs ({sum(r['kind'] == 'octree' for r in runs)} octrees), {sum(cells)} cells, "
          "64 served steps equal to the runs stepped alone", flush=True)
