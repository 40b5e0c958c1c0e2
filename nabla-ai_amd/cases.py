"""Case builders and synthetic inputs, shared by tests and bench.py.

Reference builders (core/src/flow/case.cpp:25-134) are mirrored as
`lid_cavity`, `channel` and `wall_mounted_cube`; the measured workloads are

* `taylor_green(n)`  — n^3 periodic box, L = 2*pi, Re 1600, CFL 0.5, WENO-5,
  pressure tolerance 1e-8, analytic Taylor-Green initial state;
* `cube_channel()`   — 192x64x64 channel (h = 1, 16 cells per h), uniform
  inlet U = 1 on x-, outlet on x+, no-slip floor/ceiling, periodic span, the
  staircase cube [3,4]x[0,1]x[1.5,2.5] (cell-centre rule of case.cpp:113-122),
  Re_h 500, CFL 0.7, WENO-5, 1 % seeded uniform perturbation of the initial
  interior u-faces.

The PRNG is splitmix64 with doubles from the top 53 bits, fully specified so
inputs are reproducible on any host.
"""
from __future__ import annotations

import math

import numpy as np

import nbl

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive splitmix64 outputs from `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int) -> np.ndarray:
    """Box-Muller N(0,1) from splitmix64 uniforms."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = np.maximum(u[0::2], 2.0 ** -60)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def _faces(*types):
    return [dict(type=t) for t in types]


# ---- reference builders (case.cpp:25-134) ---------------------------------
def lid_cavity(n: int, reynolds: float, lid: float = 1.0):
    faces = _faces(nbl.BC_WALL, nbl.BC_WALL, nbl.BC_WALL, nbl.BC_WALL,
                   nbl.BC_PERIODIC, nbl.BC_PERIODIC)
    faces[3]["wall_velocity"] = (lid, 0.0, 0.0)
    g = nbl.Grid(n, n, 1, 1.0 / n, 1.0 / n, 1.0 / n, faces=faces)
    cfg = nbl.flow_cfg(reynolds=reynolds, ref_length=1.0, velocity_scale=lid, cfl=0.8,
                       convection=nbl.CONV_CENTRAL)
    return g, cfg


def channel(nx: int, ny: int, reynolds: float, bulk: float = 1.0):
    ly, lx = 1.0, 6.0
    faces = _faces(nbl.BC_INLET, nbl.BC_OUTLET, nbl.BC_WALL, nbl.BC_WALL,
                   nbl.BC_PERIODIC, nbl.BC_PERIODIC)
    faces[0].update(profile=nbl.INLET_PARABOLIC, inlet_speed=bulk, inlet_axis=0)
    g = nbl.Grid(nx, ny, 1, lx / nx, ly / ny, ly / ny, faces=faces)
    cfg = nbl.flow_cfg(reynolds=reynolds, ref_length=ly, velocity_scale=bulk, cfl=0.8,
                       convection=nbl.CONV_CENTRAL)
    return g, cfg


def cube_solid(nx, ny, nz, dx, dy, dz, x0, x1, y0, y1, z0, z1):
    i = (np.arange(nx) + 0.5) * dx
    j = (np.arange(ny) + 0.5) * dy
    k = (np.arange(nz) + 0.5) * dz
    X, Y, Z = np.meshgrid(i, j, k, indexing="ij")
    s = (X >= x0) & (X <= x1) & (Y >= y0) & (Y <= y1) & (Z >= z0) & (Z <= z1)
    return np.ascontiguousarray(s.transpose(2, 1, 0)).astype(np.uint8).reshape(-1)


def wall_mounted_cube(res: int, reynolds: float):
    h = 1.0
    dx = h / res
    nx, ny, nz = round(14.0 * res), round(3.0 * res), round(6.4 * res)
    faces = _faces(nbl.BC_INLET, nbl.BC_OUTLET, nbl.BC_WALL, nbl.BC_WALL,
                   nbl.BC_PERIODIC, nbl.BC_PERIODIC)
    faces[0].update(profile=nbl.INLET_UNIFORM, inlet_speed=1.0, inlet_axis=0)
    zc = 3.2 * h
    solid = cube_solid(nx, ny, nz, dx, dx, dx, 3.0 * h, 4.0 * h, 0.0, h, zc - 0.5 * h, zc + 0.5 * h)
    g = nbl.Grid(nx, ny, nz, dx, faces=faces, solid=solid)
    cfg = nbl.flow_cfg(reynolds=reynolds, ref_length=h, velocity_scale=1.0, ref_area=h * h,
                       cfl=0.7, convection=nbl.CONV_WENO5)
    return g, cfg


# ---- measured workloads -----------------------------------------------------
def taylor_green(n: int = 64, reynolds: float = 1600.0, tol: float = 1e-8):
    L = 2.0 * math.pi
    h = L / n
    g = nbl.Grid(n, n, n, h, faces=_faces(*([nbl.BC_PERIODIC] * 6)))
    cfg = nbl.flow_cfg(reynolds=reynolds, ref_length=1.0, velocity_scale=1.0, cfl=0.5,
                       convection=nbl.CONV_WENO5, poisson_tol=tol, poisson_max_iters=5000)
    return g, cfg


def taylor_green_state(g: nbl.Grid):
    """u = sin x cos y cos z on x-faces, v = -cos x sin y cos z on y-faces, w = p = 0.
    The duplicated periodic faces (i = nx, j = ny) equal their i = 0 / j = 0 copies."""
    nx, ny, nz, h = g.nx, g.ny, g.nz, g.dx
    xf = np.arange(nx + 1) * h
    xc = (np.arange(nx) + 0.5) * h
    yf = np.arange(ny + 1) * h
    yc = (np.arange(ny) + 0.5) * h
    zc = (np.arange(nz) + 0.5) * h
    Z, Y, X = np.meshgrid(zc, yc, xf, indexing="ij")
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    u[:, :, nx] = u[:, :, 0]
    Z, Y, X = np.meshgrid(zc, yf, xc, indexing="ij")
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    v[:, ny, :] = v[:, 0, :]
    w = np.zeros(g.shapes[2])
    p = np.zeros(g.shapes[3])
    return u.reshape(-1).copy(), v.reshape(-1).copy(), w, p


def cube_channel(nx=192, ny=64, nz=64, res=16, reynolds=500.0, tol=1e-8):
    h = 1.0
    dx = h / res
    faces = _faces(nbl.BC_INLET, nbl.BC_OUTLET, nbl.BC_WALL, nbl.BC_WALL,
                   nbl.BC_PERIODIC, nbl.BC_PERIODIC)
    faces[0].update(profile=nbl.INLET_UNIFORM, inlet_speed=1.0, inlet_axis=0)
    solid = cube_solid(nx, ny, nz, dx, dx, dx, 3.0 * h, 4.0 * h, 0.0, h,
                       0.5 * nz * dx - 0.5 * h, 0.5 * nz * dx + 0.5 * h)
    g = nbl.Grid(nx, ny, nz, dx, faces=faces, solid=solid)
    cfg = nbl.flow_cfg(reynolds=reynolds, ref_length=h, velocity_scale=1.0, ref_area=h * h,
                       cfl=0.7, convection=nbl.CONV_WENO5, poisson_tol=tol,
                       poisson_max_iters=5000)
    return g, cfg


def cube_channel_state(g: nbl.Grid, seed: int = 20240601, amp: float = 0.01):
    """Initial state: u = 1 + amp*xi on interior x-faces (xi ~ U(-1,1)), inlet
    faces 1, solid-touching faces 0, v = w = p = 0."""
    nx, ny, nz = g.nx, g.ny, g.nz
    u = np.ones((nz, ny, nx + 1))
    xi = 2.0 * uniform01(seed, nz * ny * (nx - 1)).reshape(nz, ny, nx - 1) - 1.0
    u[:, :, 1:nx] += amp * xi
    s = g.solid.reshape(nz, ny, nx).astype(bool)
    touch = np.zeros((nz, ny, nx + 1), bool)
    touch[:, :, :nx] |= s
    touch[:, :, 1:] |= s
    u[touch] = 0.0
    u[:, :, nx] = u[:, :, nx - 1]
    v = np.zeros(g.shapes[1])
    w = np.zeros(g.shapes[2])
    p = np.zeros(g.shapes[3])
    return u.reshape(-1).copy(), v, w, p


def kinetic_energy_enstrophy(g: nbl.Grid, u, v, w):
    """KE = (1/2N) sum |u_c|^2 and enstrophy = (1/2N) sum |omega_c|^2 with
    cell-centred velocities (ns_solver.cpp:677-688) and periodic central
    differences for the vorticity."""
    nx, ny, nz = g.nx, g.ny, g.nz
    U = u.reshape(nz, ny, nx + 1)
    V = v.reshape(nz, ny + 1, nx)
    W = w.reshape(nz + 1, ny, nx)
    uc = 0.5 * (U[:, :, :-1] + U[:, :, 1:])
    vc = 0.5 * (V[:, :-1, :] + V[:, 1:, :])
    wc = 0.5 * (W[:-1] + W[1:]) if nz > 1 else np.zeros((nz, ny, nx))
    n = nx * ny * nz
    ke = float(np.sum(uc * uc + vc * vc + wc * wc)) / (2.0 * n)

    def d(f, axis, h):
        return (np.roll(f, -1, axis) - np.roll(f, 1, axis)) / (2.0 * h)
    ox = d(wc, 1, g.dy) - d(vc, 0, g.dz)
    oy = d(uc, 0, g.dz) - d(wc, 2, g.dx)
    oz = d(vc, 2, g.dx) - d(uc, 1, g.dy)
    ens = float(np.sum(ox * ox + oy * oy + oz * oz)) / (2.0 * n)
    return ke, ens


def poisson_rhs(g: nbl.Grid, seed: int):
    """RHS ~ N(0,1) i.i.d. on fluid cells, projected to zero mean."""
    b = normal(seed, g.cells)
    fl = g.solid == 0
    b[~fl] = 0.0
    b[fl] -= b[fl].mean()
    return b


# ---- batched solves and the served stream --------------------------------------
def batch_uniform(count: int = 48, seed: int = 4242):
    """The batched-solve problem set: n^3 unit boxes with n drawn by seed from
    32/64/96/128, all faces walls (pure Neumann, singular), i.i.d. N(0,1)
    right-hand sides projected to zero mean.  Returns (sizes, grids, rhs)."""
    u = uniform01(seed, count)
    sizes = [(32, 64, 96, 128)[int(float(x) * 4) % 4] for x in u]
    grids = [nbl.Grid(n, n, n, 1.0 / n) for n in sizes]
    rhs = []
    for q, g in enumerate(grids):
        b = normal(seed + 1 + q, g.cells)
        rhs.append(b - b.mean())
    return sizes, grids, rhs


def batch_octrees(count, seed):
    """Octrees of 0.5-2M leaves: unit cube, base level 6 (262,144 leaves), then
    seeded random boxes refined (balanced) to level 8 until the leaf count
    reaches a seeded target."""
    u = uniform01(seed, 4096)
    out, q = [], 0
    for t in range(count):
        target = int(500_000 + 1_500_000 * float(u[q]))
        q += 1
        tr = nbl.Tree((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 8)
        tr.refine_uniform(6)
        while tr.n < target:
            _k, lv, c = tr.leaves()
            f = np.zeros(lv.size, bool)
            for _ in range(3):  # three seeded boxes per balanced pass
                lo = np.array([float(u[q]), float(u[q + 1]), float(u[q + 2])]) * 0.7
                hi = lo + 0.08 + 0.22 * float(u[q + 3])
                q += 4
                f |= np.all((c >= lo) & (c <= hi), axis=1)
            f &= lv < 8
            if f.any():
                tr.refine_flagged(f)
        k, lv, _c = tr.leaves()
        out.append((k, lv))
    return out


def served_runs(nruns, seed):
    """Seeded runs of the three families, a third each: Taylor-Green n^3
    boxes, cube channels at r cells per cube side (192 r^3 cells) and
    cube-wake octrees (finest level 7 or 8, wake box 1.5-5 h long), sizes
    0.3-4M cells or leaves, Re uniform in [500, 3000].  Each run is a dict:
    name, kind ("uniform" or "octree"), cells, and either grid, cfg, state
    (u, v, w, p) or leaves (keys, levels, masks), origin, extent, max_level,
    cfg (an OctFlowCfg) and state (u, v, w)."""
    u = uniform01(seed, 4 * nruns)
    out = []
    for q in range(nruns):
        f, re, s1, s2 = (float(x) for x in u[4 * q:4 * q + 4])
        re = 500.0 + 2500.0 * re
        if f < 1.0 / 3.0:
            n = 2 * int(round((67 + (158 - 67) * s1) / 2))
            g, cfg = taylor_green(n, reynolds=re)
            out.append(dict(name=f"tgv{n}", kind="uniform", cells=g.cells, grid=g, cfg=cfg,
                            state=taylor_green_state(g)))
        elif f < 2.0 / 3.0:
            r = int(round(12 + (27 - 12) * s1))
            g, cfg = cube_channel(nx=12 * r, ny=4 * r, nz=4 * r, res=r, reynolds=re)
            out.append(dict(name=f"cube{12 * r}x{4 * r}x{4 * r}", kind="uniform", cells=g.cells, grid=g, cfg=cfg,
                            state=cube_channel_state(g, seed=seed + q)))
        else:
            level = 7 if s1 < 0.5 else 8
            wake = 1.5 + 3.5 * s2
            out.append(cube_wake_run(level, wake, re, seed + q))
    return out


def cube_wake_run(max_level, wake, reynolds, seed, precision=nbl.PREC_MIXED):
    """A cube-wake octree flow run: the cube-wake refinement rule with the wake
    box `wake` h long, finest level max_level, the cube's leaves InsideSolid;
    uniform inflow, outlet, no-slip floor and ceiling, free-slip sides, CFL
    0.7, WENO-5, FP32 V-cycle; u = 1 + 0.01 xi on fluid leaves."""
    t = nbl.build_tree(C3_ORIGIN, C3_EXTENT, max_level, 5, lambda c: c3_predicate(c, wake))
    k, l, c = t.leaves()
    t.close()
    m = c3_solid(c)
    fl = m == 0
    n = k.size
    uu = np.where(fl, 1.0 + 0.01 * (2.0 * uniform01(seed, n) - 1.0), 0.0)
    cfg = nbl.oct_flow_cfg(nu=1.0 / reynolds, cfl=0.7, inlet_speed=1.0, inlet_profile=0,
                           face=(nbl.OCT_INLET, nbl.OCT_OUTLET, nbl.OCT_WALL, nbl.OCT_WALL,
                                 nbl.OCT_SLIP, nbl.OCT_SLIP),
                           convection=1, poisson_tol=1e-8, poisson_max_iters=500, precision=precision)
    return dict(name=f"wake{max_level}_{wake:.1f}", kind="octree", cells=int(n), leaves=(k, l, m),
                origin=C3_ORIGIN, extent=C3_EXTENT, max_level=max_level, cfg=cfg,
                state=(uu, np.zeros(n), np.zeros(n)))


def admit_served(srv, r, opts=None):
    """Admit a served_runs() entry to an nbl.Server and set its state."""
    if r["kind"] == "uniform":
        rid = srv.admit(r["grid"], r["cfg"], opts)
        srv.run(rid, r["grid"]).set_state(*r["state"])
    else:
        rid = srv.admit_octree(*r["leaves"], r["origin"], r["extent"], r["max_level"], r["cfg"])
        srv.octree_flow(rid).set_state(*r["state"])
    return rid


def served_state(srv, rid, r):
    """The run's state arrays as the server holds them."""
    if r["kind"] == "uniform":
        return srv.run(rid, r["grid"]).get_state()[:4]
    return srv.octree_flow(rid).get_state()


def step_alone(ctx, r, steps, opts=None):
    """The same run stepped by itself: (reports, state arrays)."""
    if r["kind"] == "uniform":
        run = nbl.Run(ctx, r["grid"], r["cfg"], opts)
        try:
            run.set_state(*r["state"])
            reps = [run.step() for _ in range(steps)]
            return reps, run.get_state()[:4]
        finally:
            run.close()
    d = nbl.Octree(ctx, *r["leaves"], r["origin"], r["extent"], r["max_level"])
    f = nbl.OctreeFlow(d, r["cfg"])
    try:
        f.set_state(*r["state"])
        reps = [f.step() for _ in range(steps)]
        return reps, f.get_state()
    finally:
        f.close()
        d.close()


# ---- octree workload (cube wake) ---------------------------------------------
CUBE_BOX = (3.0, 4.0, 0.0, 1.0, 1.5, 2.5)
WAKE_BOX = (4.0, 10.0, 0.0, 2.0, 1.0, 3.0)
C3_ORIGIN = (0.0, 0.0, 0.0)
C3_EXTENT = (12.0, 4.0, 4.0)


def c3_predicate(centers: np.ndarray, wake: float = 6.0) -> np.ndarray:
    """Refinement rule of the cube-wake octree: centre within 1.5 h of the cube's
    bounding box, or inside the wake box [0,6h] behind the rear face, +-1 h about
    the span centre, [0,2h] above the floor (h = 1; `wake` sets the box's
    length behind the cube)."""
    x, y, z = centers[:, 0], centers[:, 1], centers[:, 2]
    x0, x1, y0, y1, z0, z1 = CUBE_BOX
    dx = np.maximum(np.maximum(x0 - x, 0.0), x - x1)
    dy = np.maximum(np.maximum(y0 - y, 0.0), y - y1)
    dz = np.maximum(np.maximum(z0 - z, 0.0), z - z1)
    near = dx * dx + dy * dy + dz * dz <= 1.5 * 1.5
    a0, _a1, b0, b1, c0, c1 = WAKE_BOX
    in_wake = (x >= a0) & (x <= a0 + wake) & (y >= b0) & (y <= b1) & (z >= c0) & (z <= c1)
    return near | in_wake


def c3_solid(centers: np.ndarray) -> np.ndarray:
    """InsideSolid mask (2) for leaves whose centre lies in the cube
    (core/src/adaptive/run.cpp:67-75), Fluid (0) elsewhere."""
    x, y, z = centers[:, 0], centers[:, 1], centers[:, 2]
    x0, x1, y0, y1, z0, z1 = CUBE_BOX
    inside = (x >= x0) & (x <= x1) & (y >= y0) & (y <= y1) & (z >= z0) & (z <= z1)
    return np.where(inside, 2, 0).astype(np.uint8)


def face_stream_sha256(keys: np.ndarray, cnt: np.ndarray, nbr: np.ndarray) -> str:
    """SHA-256 of an octree's connectivity in a layout-independent form: leaves
    in ascending key order, each with its 6 neighbour counts and the 6x4
    neighbour keys in face order (unused slots all ones).  `nbr` holds leaf
    indices into `keys` (-1 unused), in any leaf order matching `keys`."""
    import hashlib
    keys = np.asarray(keys, np.uint64)
    o = np.argsort(keys, kind="stable")
    nb = np.asarray(nbr, np.int64).reshape(len(keys), 6, 4)[o]
    nk = np.where(nb >= 0, keys[np.maximum(nb, 0)], np.uint64(0xFFFFFFFFFFFFFFFF)).astype(np.uint64)
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(keys[o]).tobytes())
    h.update(np.ascontiguousarray(np.asarray(cnt, np.uint8).reshape(len(keys), 6)[o]).tobytes())
    h.update(np.ascontiguousarray(nk).tobytes())
    return h.hexdigest()
