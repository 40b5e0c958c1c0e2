"""ctypes view of the nbl C-ABI (include/nbl_api.h) for tests and bench.py.

The product's host layer is C++20 (nabla/flow headers in this package); this
module only mirrors the POD descriptors and binds the C entry points so that
Python tests can drive the HIP library the way the reference's Catch2 tests
drive nabla_core in-process (core/tests/test_flow.cpp:103-105).

`load()` raises if the HIP library is not built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_PATH = PKG / "lib" / "libnabla_hip.so"

BC_WALL, BC_INLET, BC_OUTLET, BC_PERIODIC = 0, 1, 2, 3
INLET_UNIFORM, INLET_PARABOLIC = 0, 1
CONV_CENTRAL, CONV_WENO5 = 0, 1
PREC_FP64, PREC_MIXED, PREC_MIXED_BF16 = 0, 1, 2
COARSE_DIRECT, COARSE_CG = 0, 1
VCYCLE_AUTO, VCYCLE_PER_SWEEP, VCYCLE_TILED = 0, 1, 2


class FaceBC(C.Structure):
    _fields_ = [("type", C.c_int), ("wall_velocity", C.c_double * 3), ("profile", C.c_int),
                ("inlet_speed", C.c_double), ("inlet_axis", C.c_int)]


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("origin", C.c_double * 3),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("face", FaceBC * 6), ("solid", C.POINTER(C.c_uint8))]


class FlowCfg(C.Structure):
    _fields_ = [("reynolds", C.c_double), ("ref_length", C.c_double),
                ("velocity_scale", C.c_double), ("density", C.c_double), ("cfl", C.c_double),
                ("body_force", C.c_double * 3), ("convection", C.c_int),
                ("poisson_tol", C.c_double), ("poisson_max_iters", C.c_int),
                ("tol_continuity", C.c_double), ("tol_mass", C.c_double),
                ("ref_area", C.c_double)]


class SolverOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("pre_smooth", C.c_int), ("post_smooth", C.c_int),
                ("check_every", C.c_int), ("coarse", C.c_int), ("cells_per_thread", C.c_int),
                ("loops", C.c_int), ("rhs_planes", C.c_int), ("vcycle", C.c_int), ("launch_order", C.c_int),
                ("iteration_graph", C.c_int)]


class StepReport(C.Structure):
    _fields_ = [("step", C.c_int), ("t", C.c_double), ("dt", C.c_double), ("cfl", C.c_double),
                ("momentum_residual", C.c_double), ("continuity_residual", C.c_double),
                ("divergence_max", C.c_double), ("mass_error", C.c_double),
                ("poisson_iters", C.c_int), ("poisson_residual", C.c_double),
                ("cd", C.c_double), ("cl", C.c_double), ("cell_count", C.c_uint64),
                ("accepted", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


OCT_WALL, OCT_SLIP, OCT_INLET, OCT_OUTLET = 0, 1, 2, 3


class OctFlowCfg(C.Structure):
    _fields_ = [("nu", C.c_double), ("cfl", C.c_double), ("inlet_speed", C.c_double), ("inlet_profile", C.c_int),
                ("face", C.c_int * 6),
                ("convection", C.c_int), ("poisson_tol", C.c_double), ("poisson_max_iters", C.c_int),
                ("precision", C.c_int)]


class OctStepReport(C.Structure):
    _fields_ = [("step", C.c_int), ("t", C.c_double), ("dt", C.c_double), ("divergence_max", C.c_double),
                ("kinetic_energy", C.c_double), ("inflow", C.c_double), ("outflow", C.c_double),
                ("poisson_iters", C.c_int), ("poisson_residual", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PoissonReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("converged", C.c_int)]


P_D = C.POINTER(C.c_double)
P_U8 = C.POINTER(C.c_uint8)
P_U64 = C.POINTER(C.c_uint64)
P_I32 = C.POINTER(C.c_int32)


class SensorFields(C.Structure):
    _fields_ = [(f, P_D) for f in ("vorticity_mag", "strain_mag", "q_criterion", "grad_p_mag", "grad_u_mag",
                                   "y_plus", "tau_wall", "trunc_err")] + \
               [(f, P_U8) for f in ("near_wall", "reverse_flow", "resid_stagnation")]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(P_D)


# --------------------------------------------------------------------------
# Grid / config builders (defaults of FlowConfig, ns_solver.hpp:14-32)
# --------------------------------------------------------------------------
class Grid:
    """Host description of a uniform MAC grid; keeps the solid array alive."""

    def __init__(self, nx, ny, nz, dx, dy=None, dz=None, origin=(0.0, 0.0, 0.0),
                 faces=None, solid=None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = float(dx)
        self.dy = float(dx if dy is None else dy)
        self.dz = float(dx if dz is None else dz)
        self.origin = tuple(float(o) for o in origin)
        self.faces = faces if faces is not None else [dict(type=BC_WALL) for _ in range(6)]
        n = self.nx * self.ny * self.nz
        self.solid = (np.zeros(n, np.uint8) if solid is None
                      else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1))
        assert self.solid.size == n

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    @property
    def shapes(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1), nx * ny * nz)

    def desc(self) -> GridDesc:
        d = GridDesc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.origin[:] = self.origin
        d.dx, d.dy, d.dz = self.dx, self.dy, self.dz
        for f, spec in enumerate(self.faces):
            d.face[f].type = spec.get("type", BC_WALL)
            d.face[f].wall_velocity[:] = spec.get("wall_velocity", (0.0, 0.0, 0.0))
            d.face[f].profile = spec.get("profile", INLET_UNIFORM)
            d.face[f].inlet_speed = spec.get("inlet_speed", 0.0)
            d.face[f].inlet_axis = spec.get("inlet_axis", 0)
        d.solid = self.solid.ctypes.data_as(P_U8)
        d._keep = self.solid  # noqa: keep the buffer alive with the descriptor
        return d


def flow_cfg(**kw) -> FlowCfg:
    c = FlowCfg()
    c.reynolds, c.ref_length, c.velocity_scale, c.density, c.cfl = 100.0, 1.0, 1.0, 1.0, 0.8
    c.body_force[:] = (0.0, 0.0, 0.0)
    c.convection = CONV_CENTRAL
    c.poisson_tol, c.poisson_max_iters = 1e-9, 5000
    c.tol_continuity, c.tol_mass, c.ref_area = 1e-7, 1e-8, 0.0
    for k, v in kw.items():
        if k == "body_force":
            c.body_force[:] = v
        else:
            setattr(c, k, v)
    return c


def solver_opts(precision=PREC_FP64, pre=2, post=2, check_every=1, coarse=None, cells_per_thread=0,
                loops=False, rhs_planes=0, vcycle=0, launch_order=0, iteration_graph=0) -> SolverOpts:
    """loops: the reference's loop structure, one kernel per loop (the unfused
    FP64 baseline).  rhs_planes: planes each WENO-5 right-hand-side brick
    marches through (0 automatic).  vcycle: 0 automatic, 1 one kernel per
    sweep, 2 two kernels per level (VCYCLE_*).  launch_order: 0 automatic, 1
    plain, 2 slabs of planes per XCD, 3 bands of rows per XCD.
    iteration_graph: 0 / 1 kernel by kernel (default), 2 pairs of iterations
    replayed from a captured HIP graph."""
    o = SolverOpts()
    o.precision, o.pre_smooth, o.post_smooth, o.check_every = precision, pre, post, check_every
    o.coarse = COARSE_DIRECT if coarse is None else coarse
    o.cells_per_thread = cells_per_thread
    o.loops = 1 if loops else 0
    o.rhs_planes = rhs_planes
    o.vcycle = vcycle
    o.launch_order = launch_order
    o.iteration_graph = iteration_graph
    return o


# --------------------------------------------------------------------------
# Library loading
# --------------------------------------------------------------------------
_lib = None


def _proto(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def load(path: os.PathLike | None = None):
    """Load the HIP library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"nbl: HIP library not built ({p}); run __graft_entry__.build()")
    lib = C.CDLL(str(p))
    V = C.c_void_p
    i, d, sz = C.c_int, C.c_double, C.c_size_t
    _proto(lib, "nbl_last_error", C.c_char_p)
    _proto(lib, "nbl_version", C.c_char_p)
    _proto(lib, "nbl_flow_cfg_default", None, C.POINTER(FlowCfg))
    _proto(lib, "nbl_solver_opts_default", None, C.POINTER(SolverOpts))
    _proto(lib, "nbl_ctx_create", i, i, C.POINTER(V))
    _proto(lib, "nbl_ctx_destroy", i, V)
    _proto(lib, "nbl_ctx_bytes_in_use", i, V, P_U64)
    _proto(lib, "nbl_ctx_probe_copy", i, V, C.c_uint64, i, C.POINTER(C.c_double))
    _proto(lib, "nbl_run_create", i, V, C.POINTER(GridDesc), C.POINTER(FlowCfg),
           C.POINTER(SolverOpts), C.POINTER(V))
    _proto(lib, "nbl_run_destroy", i, V)
    _proto(lib, "nbl_run_initialize", i, V)
    _proto(lib, "nbl_run_set_state", i, V, P_D, sz, P_D, sz, P_D, sz, P_D, sz, i, d, d)
    _proto(lib, "nbl_run_get_state", i, V, P_D, P_D, P_D, P_D, C.POINTER(i), P_D, P_D)
    _proto(lib, "nbl_run_compute_dt", i, V, P_D)
    _proto(lib, "nbl_run_step", i, V, d, C.POINTER(StepReport))
    _proto(lib, "nbl_run_step_many", i, V, i, C.POINTER(StepReport))
    _proto(lib, "nbl_run_device_state", i, V, C.POINTER(P_D), C.POINTER(P_D), C.POINTER(P_D),
           C.POINTER(P_D))
    _proto(lib, "nbl_run_rhs", i, V, P_D, P_D, P_D)
    _proto(lib, "nbl_run_cell_velocity", i, V, i, P_D)
    _proto(lib, "nbl_run_last_solves", i, V, P_D, C.POINTER(i), C.POINTER(i))
    _proto(lib, "nbl_poisson_profile", i, V, i, i, P_D, P_D)
    _proto(lib, "nbl_poisson_create", i, V, C.POINTER(GridDesc), i, d, i, C.POINTER(SolverOpts),
           C.POINTER(V))
    _proto(lib, "nbl_poisson_destroy", i, V)
    _proto(lib, "nbl_poisson_solve", i, V, P_D, P_D, C.POINTER(PoissonReport))
    _proto(lib, "nbl_poisson_solve_device", i, V, C.c_void_p, C.c_void_p, C.POINTER(PoissonReport))
    _proto(lib, "nbl_poisson_solve_noalloc", i, V, P_D, P_D, C.POINTER(PoissonReport))
    _proto(lib, "nbl_poisson_precondition", i, V, P_D, P_D)
    _proto(lib, "nbl_poisson_apply", i, V, P_D, P_D)
    _proto(lib, "nbl_poisson_diagonal", i, V, P_D)
    _proto(lib, "nbl_poisson_levels", i, V, C.POINTER(i), C.POINTER(i))
    _proto(lib, "nbl_poisson_smooth_once", i, V, P_D, P_D)
    _proto(lib, "nbl_poisson_restrict", i, V, P_D, P_D)
    _proto(lib, "nbl_poisson_prolong", i, V, P_D, P_D)
    _proto(lib, "nbl_poisson_last_timing", i, V, P_D, C.POINTER(i))
    _proto(lib, "nbl_tree_create", i, P_D, P_D, i, C.POINTER(V))
    _proto(lib, "nbl_tree_destroy", i, V)
    _proto(lib, "nbl_tree_count", i, V, C.POINTER(sz))
    _proto(lib, "nbl_tree_refine_uniform", i, V, i)
    _proto(lib, "nbl_tree_refine_flagged", i, V, P_U8, sz)
    _proto(lib, "nbl_tree_is_balanced", i, V, C.POINTER(i))
    _proto(lib, "nbl_tree_leaves", i, V, P_U64, P_U8, P_D)
    _proto(lib, "nbl_octree_create", i, V, P_U64, P_U8, P_U8, sz, P_D, P_D, i, C.POINTER(V))
    _proto(lib, "nbl_octree_destroy", i, V)
    _proto(lib, "nbl_octree_count", i, V, C.POINTER(sz), C.POINTER(sz))
    _proto(lib, "nbl_octree_coef_classes", i, V, C.POINTER(i))
    _proto(lib, "nbl_octree_permutation", i, V, P_U64)
    _proto(lib, "nbl_octree_face_table", i, V, P_U8, P_I32)
    _proto(lib, "nbl_octree_apply", i, V, P_D, P_D)
    _proto(lib, "nbl_octree_diagonal", i, V, P_D)
    _proto(lib, "nbl_octree_poisson_solve", i, V, P_D, P_D, d, i, C.POINTER(PoissonReport))
    _proto(lib, "nbl_octree_transport", i, V, P_D, P_D, P_D, P_D, d, i, C.POINTER(C.c_double))
    _proto(lib, "nbl_octree_sensors", i, V, P_D, P_D, P_D, P_D, d, P_D, P_D, C.POINTER(SensorFields))
    _proto(lib, "nbl_octree_mg_solve", i, V, P_D, P_D, d, i, C.POINTER(PoissonReport))
    _proto(lib, "nbl_octree_mg_solve_from", i, V, P_D, P_D, d, i, C.POINTER(PoissonReport))
    _proto(lib, "nbl_octree_mg_levels", i, V, C.POINTER(C.c_longlong), i, C.POINTER(i),
           C.POINTER(C.c_double))
    _proto(lib, "nbl_octree_set_precision", i, V, i)
    _proto(lib, "nbl_octree_flow_cfg_default", None, C.POINTER(OctFlowCfg))
    _proto(lib, "nbl_octree_flow_create", i, V, C.POINTER(OctFlowCfg), C.POINTER(V))
    _proto(lib, "nbl_octree_flow_destroy", i, V)
    _proto(lib, "nbl_octree_flow_set_state", i, V, P_D, P_D, P_D, P_D)
    _proto(lib, "nbl_octree_flow_get_state", i, V, P_D, P_D, P_D, P_D)
    _proto(lib, "nbl_octree_flow_step", i, V, d, C.POINTER(OctStepReport))
    _proto(lib, "nbl_octree_flow_divergence", i, V, P_D)
    _proto(lib, "nbl_octree_profile", i, V, i, i, C.POINTER(C.c_double), C.POINTER(C.c_double))
    _proto(lib, "nbl_octree_last_timing", i, V, C.POINTER(C.c_double), C.POINTER(C.c_double))
    _proto(lib, "nbl_run_sweep_timing", i, V, i, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
           C.POINTER(C.c_double))
    _proto(lib, "nbl_server_create", i, V, i, C.c_uint64, C.POINTER(V))
    _proto(lib, "nbl_server_destroy", i, V)
    _proto(lib, "nbl_server_admit", i, V, V, V, V, C.POINTER(C.c_int))
    _proto(lib, "nbl_server_run", i, V, i, C.POINTER(V))
    _proto(lib, "nbl_server_admit_octree", i, V, P_U64, P_U8, P_U8, sz, P_D, P_D, i, C.POINTER(OctFlowCfg),
           C.POINTER(C.c_int))
    _proto(lib, "nbl_server_octree_flow", i, V, i, C.POINTER(V))
    _proto(lib, "nbl_server_submit", i, V, i, C.POINTER(C.c_uint64))
    _proto(lib, "nbl_server_wait", i, V, C.c_uint64, C.POINTER(StepReport), C.POINTER(C.c_double))
    _proto(lib, "nbl_server_footprint", i, V, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
    _proto(lib, "nbl_poisson_solve_batch", i, C.POINTER(V), C.POINTER(P_D), C.POINTER(P_D), i, i,
           C.POINTER(PoissonReport))
    _proto(lib, "nbl_poisson_batch_create", i, V, C.POINTER(GridDesc), i, d, i, C.POINTER(SolverOpts),
           C.POINTER(V))
    _proto(lib, "nbl_poisson_batch_size", i, V, C.POINTER(i))
    _proto(lib, "nbl_poisson_batch_solve", i, V, C.POINTER(P_D), C.POINTER(P_D), C.POINTER(PoissonReport))
    _proto(lib, "nbl_poisson_batch_solve_device", i, V, V, V, C.POINTER(PoissonReport))
    _proto(lib, "nbl_octree_solve_batch", i, C.POINTER(V), C.POINTER(P_D), C.POINTER(P_D), i, i, d, i,
           C.POINTER(PoissonReport))
    _proto(lib, "nbl_octree_mg_solve_batch", i, C.POINTER(V), C.POINTER(P_D), C.POINTER(P_D), i, i, d, i,
           C.POINTER(PoissonReport))
    _lib = lib
    return lib


class NblError(RuntimeError):
    """A non-zero status from the C-ABI; `code` is the NBL_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"nbl error {code}: {msg}")
        self.code = code


def check(rc: int):
    if rc != 0:
        msg = _lib.nbl_last_error().decode() if _lib else "?"
        raise NblError(rc, msg)


class Ctx:
    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.nbl_ctx_create(device, C.byref(self.h)))

    def bytes_in_use(self) -> int:
        v = C.c_uint64()
        check(self.lib.nbl_ctx_bytes_in_use(self.h, C.byref(v)))
        return v.value

    def probe_copy(self, nbytes: int, reps: int = 20) -> float:
        """Average ms of one device-to-device copy of nbytes."""
        ms = C.c_double()
        check(self.lib.nbl_ctx_probe_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            check(self.lib.nbl_ctx_destroy(self.h))
            self.h = C.c_void_p()


class Run:
    """nbl_run: device-resident NSSolver (ns_solver.hpp:58-89)."""

    def __init__(self, ctx: Ctx, grid: Grid, cfg: FlowCfg, opts: SolverOpts | None = None):
        self.lib, self.grid = ctx.lib, grid
        self._desc = grid.desc()
        self.h = C.c_void_p()
        check(self.lib.nbl_run_create(ctx.h, C.byref(self._desc), C.byref(cfg),
                                      C.byref(opts) if opts is not None else None,
                                      C.byref(self.h)))

    def initialize(self):
        check(self.lib.nbl_run_initialize(self.h))

    def set_state(self, u, v, w, p, step=0, t=0.0, dt_last=0.0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p)]
        check(self.lib.nbl_run_set_state(self.h, dptr(arrs[0]), arrs[0].size, dptr(arrs[1]),
                                         arrs[1].size, dptr(arrs[2]), arrs[2].size,
                                         dptr(arrs[3]), arrs[3].size, step, t, dt_last))

    def get_state(self):
        su, sv, sw, sp = self.grid.shapes
        u, v, w, p = (np.empty(s) for s in (su, sv, sw, sp))
        st, t, dt = C.c_int(), C.c_double(), C.c_double()
        check(self.lib.nbl_run_get_state(self.h, dptr(u), dptr(v), dptr(w), dptr(p),
                                         C.byref(st), C.byref(t), C.byref(dt)))
        return u, v, w, p, st.value, t.value, dt.value

    def compute_dt(self) -> float:
        v = C.c_double()
        check(self.lib.nbl_run_compute_dt(self.h, C.byref(v)))
        return v.value

    def step(self, dt: float = 0.0) -> StepReport:
        r = StepReport()
        check(self.lib.nbl_run_step(self.h, dt, C.byref(r)))
        return r

    def step_many(self, n: int):
        reps = (StepReport * n)()
        check(self.lib.nbl_run_step_many(self.h, n, reps))
        return list(reps)

    def rhs(self):
        su, sv, sw, _ = self.grid.shapes
        a, b, c = np.empty(su), np.empty(sv), np.empty(sw)
        check(self.lib.nbl_run_rhs(self.h, dptr(a), dptr(b), dptr(c)))
        return a, b, c

    def last_solves(self):
        ms = (C.c_double * 3)()
        it = (C.c_int * 3)()
        vc = (C.c_int * 3)()
        check(self.lib.nbl_run_last_solves(self.h, ms, it, vc))
        return list(ms), list(it), list(vc)

    def sweep_timing(self, enable: int = -1):
        """(avg ms per finest-level plain sweep, launches timed, bytes per launch)."""
        ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
        check(self.lib.nbl_run_sweep_timing(self.h, enable, C.byref(ms), C.byref(n), C.byref(by)))
        return ms.value, n.value, by.value

    def cell_velocity(self, comp: int):
        out = np.empty(self.grid.cells)
        check(self.lib.nbl_run_cell_velocity(self.h, comp, dptr(out)))
        return out

    def close(self):
        if self.h and not getattr(self, "_borrowed", False):
            check(self.lib.nbl_run_destroy(self.h))
        self.h = C.c_void_p()


class Poisson:
    """nbl_poisson: the PoissonSolver plugin (poisson.hpp:20-32)."""

    MG, CG = 0, 1

    def __init__(self, ctx: Ctx, grid: Grid, kind=0, tol=1e-9, max_iters=100,
                 opts: SolverOpts | None = None):
        self.lib, self.grid = ctx.lib, grid
        self._desc = grid.desc()
        self.h = C.c_void_p()
        check(self.lib.nbl_poisson_create(ctx.h, C.byref(self._desc), kind, tol, max_iters,
                                          C.byref(opts) if opts is not None else None,
                                          C.byref(self.h)))

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.grid.cells)
        rep = PoissonReport()
        check(self.lib.nbl_poisson_solve(self.h, dptr(b), dptr(x), C.byref(rep)))
        return x, rep

    def solve_noalloc(self, b):
        """nbl_poisson_solve_noalloc: the same solve with its device staging in
        the library's fixed per-GPU array (no allocation per call)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.grid.cells)
        rep = PoissonReport()
        check(self.lib.nbl_poisson_solve_noalloc(self.h, dptr(b), dptr(x), C.byref(rep)))
        return x, rep

    def solve_device(self, d_b: int, d_x: int):
        rep = PoissonReport()
        check(self.lib.nbl_poisson_solve_device(self.h, C.c_void_p(d_b), C.c_void_p(d_x),
                                                C.byref(rep)))
        return rep

    def precondition(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.grid.cells)
        check(self.lib.nbl_poisson_precondition(self.h, dptr(r), dptr(z)))
        return z

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(self.grid.cells)
        check(self.lib.nbl_poisson_apply(self.h, dptr(x), dptr(out)))
        return out

    def diagonal(self):
        out = np.empty(self.grid.cells)
        check(self.lib.nbl_poisson_diagonal(self.h, dptr(out)))
        return out

    def levels(self):
        n = C.c_int()
        dims = (C.c_int * 24)()
        check(self.lib.nbl_poisson_levels(self.h, C.byref(n), dims))
        return [tuple(dims[3 * l:3 * l + 3]) for l in range(n.value)]

    def smooth_once(self, b, x):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64)
        check(self.lib.nbl_poisson_smooth_once(self.h, dptr(b), dptr(x)))
        return x

    def restrict(self, rf):
        lv = self.levels()
        rf = np.ascontiguousarray(rf, dtype=np.float64)
        rc = np.empty(int(np.prod(lv[1])))
        check(self.lib.nbl_poisson_restrict(self.h, dptr(rf), dptr(rc)))
        return rc

    def prolong(self, ec, xf):
        ec = np.ascontiguousarray(ec, dtype=np.float64)
        xf = np.array(xf, dtype=np.float64)
        check(self.lib.nbl_poisson_prolong(self.h, dptr(ec), dptr(xf)))
        return xf

    def profile(self, kernel: int, reps: int = 100):
        ms, by = C.c_double(), C.c_double()
        check(self.lib.nbl_poisson_profile(self.h, kernel, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def last_timing(self):
        ms, vc = C.c_double(), C.c_int()
        check(self.lib.nbl_poisson_last_timing(self.h, C.byref(ms), C.byref(vc)))
        return ms.value, vc.value

    def close(self):
        if self.h:
            check(self.lib.nbl_poisson_destroy(self.h))
            self.h = C.c_void_p()


class Tree:
    """nbl_tree: host-side balanced octree construction (sorted leaves)."""

    def __init__(self, origin, extent, max_level):
        self.lib = load()
        self.origin, self.extent, self.max_level = tuple(origin), tuple(extent), max_level
        self.h = C.c_void_p()
        check(self.lib.nbl_tree_create((C.c_double * 3)(*origin), (C.c_double * 3)(*extent),
                                       max_level, C.byref(self.h)))

    @property
    def n(self):
        v = C.c_size_t()
        check(self.lib.nbl_tree_count(self.h, C.byref(v)))
        return v.value

    def refine_uniform(self, times):
        check(self.lib.nbl_tree_refine_uniform(self.h, times))

    def refine_flagged(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        check(self.lib.nbl_tree_refine_flagged(self.h, f.ctypes.data_as(P_U8), f.size))

    def balanced(self):
        v = C.c_int()
        check(self.lib.nbl_tree_is_balanced(self.h, C.byref(v)))
        return bool(v.value)

    def leaves(self):
        n = self.n
        k = np.empty(n, np.uint64)
        l = np.empty(n, np.uint8)
        c = np.empty(3 * n)
        check(self.lib.nbl_tree_leaves(self.h, k.ctypes.data_as(P_U64), l.ctypes.data_as(P_U8),
                                       dptr(c)))
        return k, l, c.reshape(n, 3)

    def close(self):
        if self.h:
            check(self.lib.nbl_tree_destroy(self.h))
            self.h = C.c_void_p()


def build_tree(origin, extent, max_level, base, predicate):
    """Refine uniformly to `base`, then refine (balanced) every leaf below
    max_level whose centre satisfies `predicate` until none is left."""
    t = Tree(origin, extent, max_level)
    t.refine_uniform(base)
    while True:
        _k, l, c = t.leaves()
        f = predicate(c) & (l < max_level)
        if not f.any():
            return t
        t.refine_flagged(f)


def oct_flow_cfg(**kw) -> OctFlowCfg:
    """nbl_octree_flow_cfg_default with keyword overrides (face: 6 NBL_OCT_* values)."""
    c = OctFlowCfg()
    load().nbl_octree_flow_cfg_default(C.byref(c))
    for k, v in kw.items():
        if k == "face":
            for d in range(6):
                c.face[d] = int(v[d])
        else:
            setattr(c, k, v)
    return c


class OctreeFlow:
    """nbl_octree_flow_*: the collocated finite-volume flow step on an
    Octree's leaves (arrays in the octree's caller order)."""

    def __init__(self, octree, cfg: OctFlowCfg):
        self.lib, self.n, self.tree = octree.lib, octree.n, octree
        self.h = C.c_void_p()
        check(self.lib.nbl_octree_flow_create(octree.h, C.byref(cfg), C.byref(self.h)))

    def set_state(self, u, v, w, p=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w)]
        pp = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        check(self.lib.nbl_octree_flow_set_state(self.h, *[dptr(a) for a in arrs],
                                                 None if pp is None else dptr(pp)))

    def get_state(self):
        out = [np.empty(self.n) for _ in range(4)]
        check(self.lib.nbl_octree_flow_get_state(self.h, *[dptr(a) for a in out]))
        return out

    def step(self, dt=0.0):
        r = OctStepReport()
        check(self.lib.nbl_octree_flow_step(self.h, dt, C.byref(r)))
        return r

    def divergence(self):
        out = np.empty(self.n)
        check(self.lib.nbl_octree_flow_divergence(self.h, dptr(out)))
        return out

    def close(self):
        if self.h and not getattr(self, "_borrowed", False):
            check(self.lib.nbl_octree_flow_destroy(self.h))
        self.h = C.c_void_p()


class PoissonBatch:
    """nbl_poisson_batch_*: `problems` independent pressure systems on one
    grid solved together (one launch per level kernel for all of them), each
    with its own flexible-CG scalars, stopping test and report."""

    def __init__(self, ctx: Ctx, grid: Grid, problems: int, tol=1e-9, max_iters=100,
                 opts: SolverOpts | None = None):
        self.lib, self.grid, self.problems = ctx.lib, grid, int(problems)
        self._desc = grid.desc()
        self.h = C.c_void_p()
        check(self.lib.nbl_poisson_batch_create(ctx.h, C.byref(self._desc), self.problems, tol, max_iters,
                                                C.byref(opts) if opts is not None else None, C.byref(self.h)))

    def size(self):
        n = C.c_int()
        check(self.lib.nbl_poisson_batch_size(self.h, C.byref(n)))
        return n.value

    def last_timing(self):
        """(device ms of the last batch solve, V-cycles it applied)."""
        ms, vc = C.c_double(), C.c_int()
        check(self.lib.nbl_poisson_last_timing(self.h, C.byref(ms), C.byref(vc)))
        return ms.value, vc.value

    def solve(self, bs):
        """bs: one right-hand side per problem (host arrays); returns (xs, reports)."""
        assert len(bs) == self.problems
        bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bs]
        xs = [np.empty(self.grid.cells) for _ in bs]
        bp = (P_D * self.problems)(*[dptr(b) for b in bs])
        xp = (P_D * self.problems)(*[dptr(x) for x in xs])
        reps = (PoissonReport * self.problems)()
        check(self.lib.nbl_poisson_batch_solve(self.h, bp, xp, reps))
        return xs, list(reps)

    def close(self):
        if self.h:
            check(self.lib.nbl_poisson_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Octree:
    """nbl_octree: Morton-sorted leaves, face tables and OctreePoisson."""

    def __init__(self, ctx: Ctx, keys, levels, masks, origin, extent, max_level):
        self.lib = ctx.lib
        self.keys = np.ascontiguousarray(keys, dtype=np.uint64)
        self.levels = np.ascontiguousarray(levels, dtype=np.uint8)
        self.masks = np.ascontiguousarray(masks, dtype=np.uint8)
        self.n = self.keys.size
        o = (C.c_double * 3)(*origin)
        e = (C.c_double * 3)(*extent)
        self.h = C.c_void_p()
        check(self.lib.nbl_octree_create(ctx.h, self.keys.ctypes.data_as(P_U64),
                                         self.levels.ctypes.data_as(P_U8),
                                         self.masks.ctypes.data_as(P_U8), self.n, o, e,
                                         max_level, C.byref(self.h)))

    def counts(self):
        n, nnz = C.c_size_t(), C.c_size_t()
        check(self.lib.nbl_octree_count(self.h, C.byref(n), C.byref(nnz)))
        return n.value, nnz.value

    def coef_classes(self):
        """Distinct coefficient values of the leaf operator when its kernels
        read them as one-byte classes (at most 256), else 0."""
        k = C.c_int()
        check(self.lib.nbl_octree_coef_classes(self.h, C.byref(k)))
        return k.value

    def permutation(self):
        perm = np.empty(self.n, np.uint64)
        check(self.lib.nbl_octree_permutation(self.h, perm.ctypes.data_as(P_U64)))
        return perm

    def face_table(self):
        cnt = np.empty(6 * self.n, np.uint8)
        nbr = np.empty(24 * self.n, np.int32)
        check(self.lib.nbl_octree_face_table(self.h, cnt.ctypes.data_as(P_U8),
                                             nbr.ctypes.data_as(P_I32)))
        return cnt.reshape(self.n, 6), nbr.reshape(self.n, 6, 4)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(self.n)
        check(self.lib.nbl_octree_apply(self.h, dptr(x), dptr(out)))
        return out

    def diagonal(self):
        out = np.empty(self.n)
        check(self.lib.nbl_octree_diagonal(self.h, dptr(out)))
        return out

    def solve(self, b, tol=1e-9, max_iters=2000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.n)
        rep = PoissonReport()
        check(self.lib.nbl_octree_poisson_solve(self.h, dptr(b), dptr(x), tol, max_iters,
                                                C.byref(rep)))
        return x, rep

    def transport(self, u, v, w, phi, dt, steps=1):
        """octreeTransportStep `steps` times: returns (phi, integral)."""
        u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
        p = np.array(phi, dtype=np.float64)
        integral = C.c_double()
        check(self.lib.nbl_octree_transport(self.h, dptr(u), dptr(v), dptr(w), dptr(p), dt, steps,
                                            C.byref(integral)))
        return p, integral.value

    def sensors(self, u, v, w, p, nu, speed_prev, resid_prev):
        """PhysicalSensorEvaluator(nu).evaluate on the leaves: (dict of the
        per-leaf sensors and flags, updated speed_prev, updated resid_prev)."""
        u, v, w, p = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p))
        sp = np.array(speed_prev, dtype=np.float64)
        rp = np.array(resid_prev, dtype=np.float64)
        res = {}
        out = SensorFields()
        for f, t in SensorFields._fields_:
            a = np.empty(self.n, np.float64 if t is P_D else np.uint8)
            res[f] = a
            setattr(out, f, a.ctypes.data_as(t))
        check(self.lib.nbl_octree_sensors(self.h, dptr(u), dptr(v), dptr(w), dptr(p), nu, dptr(sp), dptr(rp),
                                          C.byref(out)))
        return res, sp, rp

    def mg_solve(self, b, tol=1e-9, max_iters=2000, x0=None):
        """Multigrid-preconditioned flexible CG on the same system as solve()
        (from zero, or from the initial guess x0: nbl_octree_mg_solve_from)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        assert x.size == self.n
        rep = PoissonReport()
        fn = self.lib.nbl_octree_mg_solve if x0 is None else self.lib.nbl_octree_mg_solve_from
        check(fn(self.h, dptr(b), dptr(x), tol, max_iters, C.byref(rep)))
        return x, rep

    def set_precision(self, precision):
        """Arithmetic of mg_solve's V-cycle: PREC_FP64, PREC_MIXED (FP32 levels
        inside the FP64 flexible CG) or PREC_MIXED_BF16 (FP32 levels, level 0's
        pre-smoothed iterate stored as BF16)."""
        check(self.lib.nbl_octree_set_precision(self.h, int(precision)))

    def mg_levels(self, cap=32):
        """(leaves per multigrid level, device ms of the hierarchy build)."""
        counts = (C.c_longlong * cap)()
        nl, ms = C.c_int(), C.c_double()
        check(self.lib.nbl_octree_mg_levels(self.h, counts, cap, C.byref(nl), C.byref(ms)))
        return [int(counts[q]) for q in range(min(cap, nl.value))], ms.value

    def profile(self, kernel=0, reps=50):
        ms, by = C.c_double(), C.c_double()
        check(self.lib.nbl_octree_profile(self.h, kernel, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def last_timing(self):
        """(device ms of construction, device ms of the last solve)."""
        a, b = C.c_double(), C.c_double()
        check(self.lib.nbl_octree_last_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            check(self.lib.nbl_octree_destroy(self.h))
            self.h = C.c_void_p()


class Server:
    """nbl_server: device-resident request loop over concurrent runs."""

    def __init__(self, ctx: Ctx, slots: int = 4, hbm_budget: int = 64 << 30):
        self.lib = ctx.lib
        self.h = C.c_void_p()
        check(self.lib.nbl_server_create(ctx.h, slots, hbm_budget, C.byref(self.h)))
        self._keep = []
        self._octree_n = {}

    def admit(self, grid: Grid, cfg: FlowCfg, opts: SolverOpts | None = None) -> int:
        d = grid.desc()
        rid = C.c_int()
        check(self.lib.nbl_server_admit(self.h, C.byref(d), C.byref(cfg),
                                        C.byref(opts) if opts is not None else None, C.byref(rid)))
        self._keep.append(grid)
        return rid.value

    def run(self, run_id: int, grid: Grid) -> "Run":
        """A borrowed handle (owned by the server) for set/get state."""
        h = C.c_void_p()
        check(self.lib.nbl_server_run(self.h, run_id, C.byref(h)))
        r = Run.__new__(Run)
        r.lib, r.grid, r.h, r._borrowed = self.lib, grid, h, True
        return r

    def admit_octree(self, keys, levels, masks, origin, extent, max_level, cfg: OctFlowCfg) -> int:
        """An octree flow run (nbl_server_admit_octree); leaves in caller order."""
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        lv = np.ascontiguousarray(levels, dtype=np.uint8)
        m = np.ascontiguousarray(masks, dtype=np.uint8)
        o = (C.c_double * 3)(*origin)
        e = (C.c_double * 3)(*extent)
        rid = C.c_int()
        check(self.lib.nbl_server_admit_octree(self.h, k.ctypes.data_as(P_U64), lv.ctypes.data_as(P_U8),
                                               m.ctypes.data_as(P_U8), k.size, o, e, max_level, C.byref(cfg),
                                               C.byref(rid)))
        self._octree_n[rid.value] = k.size
        return rid.value

    def octree_flow(self, run_id: int) -> "OctreeFlow":
        """A borrowed flow handle (owned by the server) for set/get state."""
        h = C.c_void_p()
        check(self.lib.nbl_server_octree_flow(self.h, run_id, C.byref(h)))
        f = OctreeFlow.__new__(OctreeFlow)
        f.lib, f.n, f.tree, f.h, f._borrowed = self.lib, self._octree_n[run_id], None, h, True
        return f

    def submit(self, run_id: int) -> int:
        t = C.c_uint64()
        check(self.lib.nbl_server_submit(self.h, run_id, C.byref(t)))
        return t.value

    def wait(self, ticket: int):
        rep = StepReport()
        ms = C.c_double()
        check(self.lib.nbl_server_wait(self.h, ticket, C.byref(rep), C.byref(ms)))
        return rep, ms.value

    def footprint(self):
        u, b = C.c_uint64(), C.c_uint64()
        check(self.lib.nbl_server_footprint(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def close(self):
        if self.h:
            check(self.lib.nbl_server_destroy(self.h))
            self.h = C.c_void_p()


def solve_batch(solvers, rhs, concurrency=4):
    """nbl_poisson_solve_batch over Poisson objects and host vectors."""
    n = len(solvers)
    hs = (C.c_void_p * n)(*[p.h for p in solvers])
    bs = [np.ascontiguousarray(b, dtype=np.float64) for b in rhs]
    xs = [np.empty(p.grid.cells) for p in solvers]
    bp = (P_D * n)(*[dptr(b) for b in bs])
    xp = (P_D * n)(*[dptr(x) for x in xs])
    reps = (PoissonReport * n)()
    check(load().nbl_poisson_solve_batch(hs, bp, xp, n, concurrency, reps))
    return xs, list(reps)


def octree_solve_batch(octs, rhs, tol=1e-9, max_iters=2000, concurrency=4, mg=False):
    """nbl_octree_solve_batch (mg: nbl_octree_mg_solve_batch) over Octree
    objects and host vectors."""
    n = len(octs)
    hs = (C.c_void_p * n)(*[o.h for o in octs])
    bs = [np.ascontiguousarray(b, dtype=np.float64) for b in rhs]
    xs = [np.empty(o.n) for o in octs]
    bp = (P_D * n)(*[dptr(b) for b in bs])
    xp = (P_D * n)(*[dptr(x) for x in xs])
    reps = (PoissonReport * n)()
    fn = load().nbl_octree_mg_solve_batch if mg else load().nbl_octree_solve_batch
    check(fn(hs, bp, xp, n, concurrency, tol, max_iters, reps))
    return xs, list(reps)
