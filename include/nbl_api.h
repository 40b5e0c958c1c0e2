/*
 * nbl_api.h — C-ABI of the MI355X (gfx950) nabla_solve hot path.
 *
 * Plain C: POD descriptors, pointers and sizes only. No C++ exception crosses
 * this boundary; every entry point returns an int status (NBL_OK == 0, < 0 on
 * error) and the message of the last failure on the calling thread is
 * available from nbl_last_error().  The C++20 host layer (the nabla/flow headers in
 * the package) wraps these calls and rethrows std::runtime_error, which is the
 * reference's error convention (core/src/flow/ns_solver.cpp:830,
 * core/src/flow/case.cpp:303-314).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   nbl_grid_desc / nbl_face_bc      <- nabla::flow::UniformGrid, FaceBC
 *                                       (core/include/nabla/flow/grid.hpp:17-73)
 *   nbl_flow_cfg / nbl_step_report   <- nabla::flow::FlowConfig, StepReport
 *                                       (core/include/nabla/flow/ns_solver.hpp:14-49)
 *   nbl_run_*                        <- nabla::flow::NSSolver
 *                                       (core/include/nabla/flow/ns_solver.hpp:58-89)
 *   nbl_poisson_*                    <- nabla::flow::PoissonSolver / MgPoisson /
 *                                       CgPoisson (core/include/nabla/flow/poisson.hpp:20-89)
 *   nbl_octree_*                     <- nabla::mesh::Octree connectivity and
 *                                       nabla::flow::OctreePoisson
 *                                       (core/include/nabla/mesh/octree.hpp:104-178,
 *                                        core/include/nabla/flow/octree_poisson.hpp:29-55)
 *
 * Threading: one nbl_run / nbl_poisson / nbl_octree is used by one host thread
 * at a time; each owns its own HIP stream.  A context may be shared.
 */
#ifndef NBL_API_H
#define NBL_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBL_OK 0
#define NBL_ERR_INVALID (-1)   /* bad argument / size mismatch            */
#define NBL_ERR_DEVICE (-2)    /* HIP runtime or kernel launch failure     */
#define NBL_ERR_STATE (-3)     /* call not valid in the object's state     */
#define NBL_ERR_CAPACITY (-4)  /* request exceeds a fixed capacity         */

/* Boundary kinds, same order as nabla::flow::BoundaryType (grid.hpp:17). */
enum { NBL_BC_WALL = 0, NBL_BC_INLET = 1, NBL_BC_OUTLET = 2, NBL_BC_PERIODIC = 3 };
/* Inlet profiles, same order as nabla::flow::InletProfile (grid.hpp:18). */
enum { NBL_INLET_UNIFORM = 0, NBL_INLET_PARABOLIC = 1 };
/* Convection schemes, same order as nabla::flow::ConvectionScheme (ns_solver.hpp:12). */
enum { NBL_CONV_CENTRAL = 0, NBL_CONV_WENO5 = 1 };
/* Pressure-solve precision. */
enum {
  NBL_PREC_FP64 = 0,  /* plain: every level of the V-cycle in FP64                  */
  NBL_PREC_MIXED = 1, /* FP32 V-cycle inside FP64 flexible CG (FP64 residual/dots)   */
  NBL_PREC_MIXED_BF16 = 2 /* as MIXED, the finest level's smoother scratch (the iterate
                             after the pre-sweeps, passed from the cycle's down to its
                             up kernel) stored as BF16; arithmetic stays FP32.  Needs
                             the two-kernels-per-level cycle (see vcycle), else
                             NBL_ERR_INVALID                                        */
};
/* Coarsest-level solve of the V-cycle.  The reference runs CgPoisson(1e-6, 500)
   there ("near-exact", multigrid.cpp:178-183). */
enum {
  NBL_COARSE_DIRECT = 0, /* exact: the level's inverse, formed once on the device
                            (projected onto the fluid mean-zero space when the
                            system is singular); falls back to CG when it cannot
                            be formed or the level exceeds 4096 cells          */
  NBL_COARSE_CG = 1      /* the reference's Jacobi-PCG to 1e-6, <= 500 iterations */
};

/* One domain face (grid.hpp:20-26). */
typedef struct nbl_face_bc {
  int type;                 /* NBL_BC_*                                  */
  double wall_velocity[3];  /* Wall: prescribed velocity (lid)            */
  int profile;              /* NBL_INLET_*                               */
  double inlet_speed;       /* bulk inlet speed                          */
  int inlet_axis;           /* velocity component the inlet drives        */
} nbl_face_bc;

/* Uniform MAC grid (grid.hpp:32-73).  Faces: 0=-x 1=+x 2=-y 3=+y 4=-z 5=+z.
 * `solid` is a host array of nx*ny*nz bytes in the compact order
 * c = i + nx*(j + ny*k) (grid.hpp:48-53); NULL means no solid cells. */
typedef struct nbl_grid_desc {
  int nx, ny, nz;
  double origin[3];
  double dx, dy, dz;
  nbl_face_bc face[6];
  const uint8_t* solid;
} nbl_grid_desc;

/* FlowConfig (ns_solver.hpp:14-32).  Defaults: nbl_flow_cfg_default(). */
typedef struct nbl_flow_cfg {
  double reynolds;
  double ref_length;
  double velocity_scale;
  double density;
  double cfl;
  double body_force[3];
  int convection;           /* NBL_CONV_* */
  double poisson_tol;
  int poisson_max_iters;
  double tol_continuity;
  double tol_mass;
  double ref_area;          /* 0 => ref_length^2 */
} nbl_flow_cfg;

/* Options of the GPU pressure solver that have no reference counterpart. */
typedef struct nbl_solver_opts {
  int precision;            /* NBL_PREC_*                                   */
  int pre_smooth;           /* damped-Jacobi sweeps before restriction (2)  */
  int post_smooth;          /* sweeps after prolongation (2)                */
  int check_every;          /* host polls the device convergence flag after
                               this many flexible-CG iterations (>=1)       */
  int coarse;               /* NBL_COARSE_* (default DIRECT)                */
  int cells_per_thread;     /* consecutive cells along x per thread in the
                               level kernels: 0 automatic, 1, 2 or 4 (a level
                               whose nx it does not divide uses 1); results
                               agree to rounding across choices            */
  int loops;                /* 1: the reference's loop structure, one kernel
                               per loop of multigrid.cpp / poisson.cpp with
                               its CgPoisson(1e-6, 500) coarse solve (FP64
                               only): the unfused baseline the fused path is
                               measured against; 0 (default): fused         */
  int rhs_planes;           /* WENO-5 right-hand side: planes along z each
                               64 x 4 face brick marches through, 0
                               automatic or 1..16; results are bit-identical
                               across choices                              */
  int vcycle;               /* V(2,2) cycle form.  1: one kernel per sweep
                               (and per restriction).  2: two kernels per
                               level (both pre-sweeps, the residual and the
                               restriction in one; both post-sweeps in one;
                               LDS tiles recomputing their halos) wherever
                               it applies: at least three levels, direct
                               coarsest solve, pre = post = 2, a finest
                               level of at most 4096 tiles (2^21 cells in
                               3-D).  0 (default): 2 where the BF16
                               smoother scratch needs it, 1 otherwise
                               (measured faster)                           */
  int launch_order;         /* order in which the level kernels' workgroups
                               walk a level (workgroups go to the 8 XCDs
                               round-robin, each XCD has its own L2): 1 plain
                               launch order, 2 slabs of planes per XCD, 3
                               bands of rows per XCD, 0 (default) 2 on levels
                               below 8M cells with at least 8 planes, 3 from
                               8M cells, 1 otherwise.  Per-cell results are
                               bit-identical across orders; sums differ only
                               through their partial-sum tree              */
  int iteration_graph;      /* flexible-CG iterations after the first: 2
                               replayed in pairs from a HIP graph captured on
                               the first solve (one launch per pair instead
                               of one per kernel); 0 (default) or 1 launched
                               kernel by kernel (measured 1-2% faster on the
                               channel and Taylor-Green steps).  Bit-identical
                               results either way                          */
} nbl_solver_opts;

/* StepReport (ns_solver.hpp:34-49). */
typedef struct nbl_step_report {
  int step;
  double t;
  double dt;
  double cfl;
  double momentum_residual;
  double continuity_residual;
  double divergence_max;
  double mass_error;
  int poisson_iters;
  double poisson_residual;
  double cd;
  double cl;
  uint64_t cell_count;
  int accepted;
} nbl_step_report;

/* PoissonSolver::Report (poisson.hpp:25-29). */
typedef struct nbl_poisson_report {
  int iterations;
  double residual;
  int converged;
} nbl_poisson_report;

typedef struct nbl_ctx nbl_ctx;
typedef struct nbl_run nbl_run;
typedef struct nbl_poisson nbl_poisson;
typedef struct nbl_octree nbl_octree;
typedef struct nbl_octree_flow nbl_octree_flow;
typedef struct nbl_server nbl_server;
typedef struct nbl_tree nbl_tree;

/* ---- defaults and errors ------------------------------------------------ */
void nbl_flow_cfg_default(nbl_flow_cfg* cfg);        /* FlowConfig{} values        */
void nbl_solver_opts_default(nbl_solver_opts* opts); /* FP64, V(2,2), check 1, direct coarse */
const char* nbl_last_error(void);
const char* nbl_version(void);

/* ---- device context ------------------------------------------------------ */
int nbl_ctx_create(int device, nbl_ctx** out);
int nbl_ctx_destroy(nbl_ctx* ctx);
/* HBM bytes currently held by runs / solvers / octrees of this context. */
int nbl_ctx_bytes_in_use(nbl_ctx* ctx, uint64_t* bytes);
/* No counterpart in the reference: measurement.  Average time (ms) of one
 * device-to-device copy of `bytes` (a plain copy kernel of known traffic, used
 * to calibrate hardware byte counters). */
int nbl_ctx_probe_copy(nbl_ctx* ctx, uint64_t bytes, int reps, double* ms);

/* ---- MAC Navier-Stokes run (NSSolver, ns_solver.hpp:58-89) ----------------
 * State stays in HBM between calls.  Host arrays use the reference's face
 * layouts: u (nx+1)*ny*nz, v nx*(ny+1)*nz, w nx*ny*(nz+1), p nx*ny*nz
 * (ns_solver.hpp:93-110).  Passing NULL for `opts` selects the defaults.     */
int nbl_run_create(nbl_ctx* ctx, const nbl_grid_desc* grid, const nbl_flow_cfg* cfg,
                   const nbl_solver_opts* opts, nbl_run** out);
int nbl_run_destroy(nbl_run* run);
/* NSSolver::initialize (ns_solver.cpp:95-112). */
int nbl_run_initialize(nbl_run* run);
/* NSSolver::setState (ns_solver.cpp:826-839); sizes are checked. */
int nbl_run_set_state(nbl_run* run, const double* u, size_t nu, const double* v, size_t nv,
                      const double* w, size_t nw, const double* p, size_t np, int step,
                      double t, double dt_last);
/* uFace/vFace/wFace/pressure + stepIndex/time/dtLast (ns_solver.hpp:142-160). Any
 * pointer may be NULL. */
int nbl_run_get_state(nbl_run* run, double* u, double* v, double* w, double* p, int* step,
                      double* t, double* dt_last);
/* NSSolver::computeDt (ns_solver.cpp:581-603). */
int nbl_run_compute_dt(nbl_run* run, double* dt);
/* NSSolver::step(dt) (ns_solver.cpp:767-824); dt <= 0 selects step() with the
 * CFL dt (ns_solver.cpp:765), computed on the device. */
int nbl_run_step(nbl_run* run, double dt, nbl_step_report* rep);
/* Advance `nsteps` CFL steps without a host round trip per step for the
 * report; `reps` (may be NULL) receives nsteps reports. */
int nbl_run_step_many(nbl_run* run, int nsteps, nbl_step_report* reps);
/* Device pointers of the resident state (u, v, w, p; reference layouts). */
int nbl_run_device_state(nbl_run* run, double** u, double** v, double** w, double** p);
/* Right-hand side of the momentum equation for the current state, as
 * NSSolver::computeRhs (ns_solver.cpp:311-473) leaves it: ru, rv, rw in the
 * face layouts, zero on boundary-plane and solid-touching faces.  The state's
 * boundary velocities are first enforced (ns_solver.cpp:775).              */
int nbl_run_rhs(nbl_run* run, double* ru, double* rv, double* rw);
/* NSSolver::cellVelocity(component) (ns_solver.cpp:690-701). */
int nbl_run_cell_velocity(nbl_run* run, int component, double* out);

/* A pressure solver that runs outside the library (the reference's
 * PoissonSolver seam, ns_solver.hpp:60-61: NSSolver takes any implementation).
 * After it is set, every projection of this run (ns_solver.cpp:501) copies its
 * right-hand side b (nx*ny*nz, compact order) to the host, calls
 * solve(user, b, x, n, rep), which must assign x (length n) and may fill rep
 * (iterations, residual, converged), and copies x back; a non-zero return
 * fails the step with NBL_ERR_STATE.  NULL restores the device MG-PCG solve. */
typedef int (*nbl_host_poisson_fn)(void* user, const double* b, double* x, size_t n, nbl_poisson_report* rep);
int nbl_run_set_pressure_solver(nbl_run* run, nbl_host_poisson_fn solve, void* user);

/* Per-stage pressure solves of the last step: device milliseconds, flexible-CG
 * iterations and V-cycles (any pointer may be NULL; arrays of 3).  No
 * reference counterpart (the reference keeps no timings). */
int nbl_run_last_solves(nbl_run* run, double* ms, int* iters, int* vcycles);

/* ---- pressure Poisson plugin (PoissonSolver, poisson.hpp:20-32) ----------
 * kind: 0 = MG-PCG "mg-pcg" (MgPoisson, multigrid.cpp:206-355),
 *       1 = Jacobi-PCG "cg-jacobi" (CgPoisson, poisson.cpp:136-252).
 * Solve A x = b (A = -Laplacian with the grid's closure).  x is assigned
 * (length nx*ny*nz).  Non-convergence is reported, not an error.            */
int nbl_poisson_create(nbl_ctx* ctx, const nbl_grid_desc* grid, int kind, double tol,
                       int max_iters, const nbl_solver_opts* opts, nbl_poisson** out);
int nbl_poisson_destroy(nbl_poisson* ps);
/* Host buffers. */
int nbl_poisson_solve(nbl_poisson* ps, const double* b, double* x, nbl_poisson_report* rep);
/* nbl_poisson_solve without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc): the device copies of b and x live in a fixed array
 * that every GPU the library runs on holds until the process ends, instead of
 * two allocations per call.  Same arguments, results and report as
 * nbl_poisson_solve; grids above 2^21 cells return NBL_ERR_CAPACITY.  All
 * calls on one GPU share that array: calls that overlap in time on one GPU
 * race (serialise them). */
int nbl_poisson_solve_noalloc(nbl_poisson* ps, const double* b, double* x, nbl_poisson_report* rep);
/* Device buffers (HBM resident, length nx*ny*nz, 16-byte aligned as any
 * hipMalloc allocation is). */
int nbl_poisson_solve_device(nbl_poisson* ps, const double* d_b, double* d_x,
                             nbl_poisson_report* rep);
/* CgPoisson::applyA (poisson.cpp:102-134), host buffers. */
/* One application of the MG-PCG preconditioner, z = Minv r: a V(pre,post)
 * cycle from zero followed by the null-space projection on singular systems
 * (the Minv of MgPoisson::solve, multigrid.cpp:283-289).  Host vectors. */
int nbl_poisson_precondition(nbl_poisson* ps, const double* r, double* z);
int nbl_poisson_apply(nbl_poisson* ps, const double* x, double* out);
/* CgPoisson::buildDiagonal (poisson.cpp:78-91) and the number of multigrid
 * levels with each level's dims (dims: 3*nlevels ints, may be NULL). */
int nbl_poisson_diagonal(nbl_poisson* ps, double* diag);
int nbl_poisson_levels(nbl_poisson* ps, int* nlevels, int* dims);
/* One level-0 damped-Jacobi sweep x <- x + 0.8 (b - A x)/diag on fluid cells,
 * without the null-space projection (multigrid.cpp:102-109). */
int nbl_poisson_smooth_once(nbl_poisson* ps, const double* b, double* x);
/* Level-0 -> level-1 full-weighting restriction of r (multigrid.cpp:127-150,
 * without the projection) and level-1 -> level-0 prolongation x += P e
 * (multigrid.cpp:158-173). */
int nbl_poisson_restrict(nbl_poisson* ps, const double* rf, double* rc);
int nbl_poisson_prolong(nbl_poisson* ps, const double* ec, double* xf);
/* Timings of the last solve: milliseconds of device time and the number of
 * preconditioner applications. */
int nbl_poisson_last_timing(nbl_poisson* ps, double* ms, int* vcycles);
/* Measurement hook (no reference counterpart): launch one kernel of the solve
 * `reps` times back to back on the solver's stream between two HIP events and
 * return the average duration of one launch (ms) and its algorithmic HBM
 * bytes (every array the kernel must touch counted once):
 *   kernel 0: finest-level damped-Jacobi sweep as the V-cycle runs it
 *             (reads x, b, 2-byte code; writes x),
 *   kernel 1: finest-level A p with the p.Ap fold (reads p, code; writes Ap),
 *   kernel 2: finest-level residual + restriction (reads b, x, code; writes
 *             the coarse residual),
 *   kernel 3: one coarsest-level solve (direct: reads the n^2 inverse, b and
 *             code, writes x; CG: counted as b, x and code only),
 *   kernel 4: finest-level down kernel of the two-kernels-per-level cycle
 *             from its right-hand side (both pre-sweeps, the residual and
 *             the restriction: reads b and code; writes the pre-smoothed
 *             iterate, BF16 in that form, and the coarse residual, reading
 *             the coarse code),
 *   kernel 5: finest-level A p of flexible-CG iterations after the first,
 *             with p = z + beta p_old formed at every stencil point and the
 *             p.Ap fold (reads z, p_old, code; writes p, Ap). */
int nbl_poisson_profile(nbl_poisson* ps, int kernel, int reps, double* ms, double* bytes);

/* ---- octree construction (host; Octree::refineUniform / refineByMorton /
 * isBalanced, octree.cpp:325-388,544-560) ------------------------------------
 * A leaf set under construction, kept sorted by Morton key.  Leaf indices in
 * the calls below refer to that sorted order. */
int nbl_tree_create(const double origin[3], const double extent[3], int max_level, nbl_tree** out);
int nbl_tree_destroy(nbl_tree* t);
int nbl_tree_count(nbl_tree* t, size_t* n);
int nbl_tree_refine_uniform(nbl_tree* t, int times);
/* Split every flagged leaf (flags: n bytes, n == count), refining coarser face
 * neighbours first so the tree stays 2:1 balanced, as refineByMorton does. */
int nbl_tree_refine_flagged(nbl_tree* t, const uint8_t* flags, size_t n);
int nbl_tree_is_balanced(nbl_tree* t, int* balanced);
/* keys, levels (n each) and cell centres (3n, octree.cpp:84-95); any may be NULL. */
int nbl_tree_leaves(nbl_tree* t, uint64_t* keys, uint8_t* levels, double* centers);

/* ---- octree (Octree connectivity + OctreePoisson) ----------------------
 * Leaves are given in any order (keys = Morton anchors on the 2^max_level grid,
 * morton.hpp:40-42); they are Morton-sorted on the device.  perm[i] (sorted
 * position -> input index) is returned by nbl_octree_permutation.           */
int nbl_octree_create(nbl_ctx* ctx, const uint64_t* keys, const uint8_t* levels,
                      const uint8_t* masks, size_t n, const double origin[3],
                      const double extent[3], int max_level, nbl_octree** out);
int nbl_octree_destroy(nbl_octree* oc);
int nbl_octree_count(nbl_octree* oc, size_t* n, size_t* nnz);
/* Storage of the leaf operator's coefficients (no reference counterpart): the
 * number of distinct coefficient values when they are at most 256 and the
 * operator kernels read a one-byte class per coupling and a table of the
 * values (the same bits), or 0 when they read one coefficient per coupling. */
int nbl_octree_coef_classes(nbl_octree* oc, int* ncls);
int nbl_octree_permutation(nbl_octree* oc, uint64_t* perm);
/* Face table in sorted order, as Octree::faceNeighbors (octree.cpp:229-283)
 * returns it: for leaf i and direction d (0..5), count[6*i+d] in {0,1,..,4}
 * and nbr[24*i + 4*d + m] = sorted index of the m-th neighbour (-1 unused). */
int nbl_octree_face_table(nbl_octree* oc, uint8_t* count, int32_t* nbr);
/* OctreePoisson::applyA / buildDiagonal (octree_poisson.cpp:64-105); arrays
 * are in the caller's input order. */
int nbl_octree_apply(nbl_octree* oc, const double* x, double* out);
int nbl_octree_diagonal(nbl_octree* oc, double* diag);
/* OctreePoisson::solve (octree_poisson.cpp:107-197): Jacobi-PCG, singular. */
int nbl_octree_poisson_solve(nbl_octree* oc, const double* b, double* x, double tol,
                             int max_iters, nbl_poisson_report* rep);
/* octreeTransportStep (adaptive/transport.cpp:24-60): first-order
 * conservative upwind update of the leaf scalar phi with cell-centred
 * velocities u, v, w, closed domain boundaries, applied `steps` times with the
 * fields resident on the device; phi in/out and the velocities are in the
 * caller's leaf order.  *integral = sum_i phi_i vol_i after the last step
 * (octreeFieldIntegral, transport.cpp:9-17). */
int nbl_octree_transport(nbl_octree* oc, const double* u, const double* v, const double* w, double* phi,
                         double dt, int steps, double* integral);

/* Results of PhysicalSensorEvaluator::evaluate (adaptive/adaptive_core.cpp:73-176,
 * CellSensors in adaptive.hpp:51-63), one entry per leaf in the caller's
 * order; any pointer may be NULL. */
typedef struct nbl_sensor_fields {
  double* vorticity_mag;
  double* strain_mag;
  double* q_criterion;
  double* grad_p_mag;
  double* grad_u_mag;
  double* y_plus;        /* -1 on cells that are not near a wall */
  double* tau_wall;
  double* trunc_err;
  uint8_t* near_wall;
  uint8_t* reverse_flow;
  uint8_t* resid_stagnation;
} nbl_sensor_fields;

/* PhysicalSensorEvaluator(nu).evaluate on the leaves: gradients of u, v, w, p
 * across level jumps (Octree::gradient, octree.cpp:491-539) from the device
 * face table, vorticity, strain, Q criterion, near-wall shear and y+, the
 * truncation estimate and the residual-stagnation flag.  u, v, w, p are
 * cell-centred leaf fields; speed_prev / resid_prev are the evaluator's
 * carried per-leaf state ("speed_prev", "resid_prev"), read and updated. */
int nbl_octree_sensors(nbl_octree* oc, const double* u, const double* v, const double* w, const double* p,
                       double nu, double* speed_prev, double* resid_prev, const nbl_sensor_fields* out);

/* Octree multigrid-preconditioned flexible CG on the OctreePoisson system.
 * No counterpart in the reference, which solves it by Jacobi-PCG only
 * (octree_poisson.cpp:107-197) and names octree levels as the intended
 * multigrid levels (docs/adr/0007-*.md).  The iteration is MgPoisson::solve's
 * (multigrid.cpp:267-355, singular form) with one V(2,2) cycle over the
 * tree's coarsenings: level l+1 merges the finest leaves of level l into
 * their parents, the operator is rediscretised with the same face
 * coefficients, residuals restrict by summation, corrections prolong by
 * injection, and the coarsest level (<= 2048 leaves) is solved exactly.
 * Same arguments and report as nbl_octree_poisson_solve; the hierarchy is
 * built on the first call. */
int nbl_octree_mg_solve(nbl_octree* oc, const double* b, double* x, double tol, int max_iters,
                        nbl_poisson_report* rep);
/* nbl_octree_mg_solve started from the guess x holds on entry instead of zero
 * (a previous step's pressure): r0 = b - A x0, the same stopping test
 * |r| <= tol |b|, no iteration when x0 already meets it; x receives the
 * solution, mean-free on the fluid leaves.  No reference counterpart: the
 * reference's solves start from zero (octree_poisson.cpp:110). */
int nbl_octree_mg_solve_from(nbl_octree* oc, const double* b, double* x, double tol, int max_iters,
                             nbl_poisson_report* rep);
/* Hierarchy of nbl_octree_mg_solve (built if needed): level count, leaves per
 * level (up to cap entries) and the device time of its construction. */
int nbl_octree_mg_levels(nbl_octree* oc, long long* counts, int cap, int* levels, double* build_ms);
/* Arithmetic of nbl_octree_mg_solve's V-cycle (NBL_PREC_*; FP64 by default):
 * NBL_PREC_MIXED runs every level's sweeps, restriction, prolongation and
 * coarsest solve in FP32 (FP32 copies of the level operators) inside the FP64
 * flexible CG, whose residuals, dot products, means and stopping test stay
 * FP64, so the FP64 tolerance is still reached; NBL_PREC_MIXED_BF16 does the
 * same with level 0's pre-smoothed iterate stored as BF16 between the sweep
 * that writes it and the restriction and post-sweep that read it.  No
 * counterpart in the reference (its octree system is solved by FP64
 * Jacobi-PCG only, octree_poisson.cpp:107-197).  Applies from the next solve;
 * NBL_ERR_INVALID for other values. */
int nbl_octree_set_precision(nbl_octree* oc, int precision);

/* ---- octree flow step (no reference implementation: the collocated
 * finite-volume solve the reference's design record plans on the octree
 * leaves, docs/adr/0008-octree-native-flow-solve.md:35-86, over its
 * OctreePoisson operator, octree_poisson.cpp:64-105) -------------------------
 * Cell-centred u, v, w, p on the leaves; face-normal velocities on the face
 * table's faces.  A step is SSP-RK3 (the MAC solver's stages, ns_solver.cpp:
 * 773-795) with a projection after each stage: provisional face velocities
 * 0.5 (u_P + u_N).n, A p = -(sum_f U_f A_f)/dt with the OctreePoisson
 * operator (area_f / dist_f couplings, Neumann at the domain boundary and at
 * non-fluid leaves; the octree multigrid CG, nbl_octree_mg_solve), then
 * U_f -= dt (p_N - p_P)/dist_f and u_P -= dt (Gauss gradient of p)_P.  The
 * momentum right-hand side is -(1/V) sum_f U_f A_f phi_f + (nu/V) sum_f A_f
 * (phi_N - phi_P)/dist_f, phi_f WENO-5 where the face has a five-leaf line
 * of one level through it and first-order upwind otherwise (at 2:1 jumps),
 * every face term formed identically from both sides (conservative across
 * level jumps).  Domain faces: NBL_OCT_WALL no-slip, NBL_OCT_SLIP free-slip,
 * NBL_OCT_INLET uniform inflow of inlet_speed along the inward normal,
 * NBL_OCT_OUTLET zero-gradient outflow scaled to the inflow's volume flux
 * (so the pressure system stays compatible).  Non-fluid leaves are no-slip
 * walls with zero velocity. */
enum { NBL_OCT_WALL = 0, NBL_OCT_SLIP = 1, NBL_OCT_INLET = 2, NBL_OCT_OUTLET = 3 };
typedef struct nbl_octree_flow_cfg {
  double nu;               /* kinematic viscosity                                  */
  double cfl;              /* dt = cfl / max over fluid leaves of |u|/hx + |v|/hy +
                              |w|/hz + 2 nu (1/hx^2 + 1/hy^2 + 1/hz^2) (the MAC
                              solver's rule, ns_solver.cpp:581-603, per leaf)      */
  double inlet_speed;      /* mean inflow speed                                    */
  int inlet_profile;       /* 0 uniform; 1 plane Poiseuille across the next axis
                              (x inlet: y): 6 s (1 - s), s the leaf centre's
                              fraction of the domain along it, times inlet_speed */
  int face[6];             /* -x, +x, -y, +y, -z, +z: NBL_OCT_*                    */
  int convection;          /* 0 first-order upwind, 1 WENO-5 away from level jumps */
  double poisson_tol;      /* relative residual of each projection's solve         */
  int poisson_max_iters;
  int precision;           /* the multigrid cycle's arithmetic, NBL_PREC_*         */
} nbl_octree_flow_cfg;
typedef struct nbl_octree_step_report {
  int step;
  double t, dt;
  double divergence_max;   /* max over fluid leaves of |sum_f U_f A_f| / V after
                              the step's last projection                          */
  double kinetic_energy;   /* sum over fluid leaves of 0.5 |u|^2 V                 */
  double inflow, outflow;  /* boundary volume fluxes of the last projection        */
  int poisson_iters;       /* the three projections' iterations                    */
  double poisson_residual; /* largest of their final relative residuals            */
  double solve_ms;         /* device ms of the three pressure solves               */
} nbl_octree_step_report;
void nbl_octree_flow_cfg_default(nbl_octree_flow_cfg* cfg);
/* The flow keeps using the octree: destroy the flow first. */
int nbl_octree_flow_create(nbl_octree* oc, const nbl_octree_flow_cfg* cfg, nbl_octree_flow** out);
int nbl_octree_flow_destroy(nbl_octree_flow* f);
/* Caller-order leaf arrays (n each); p may be NULL (zero).  Face velocities
 * start as 0.5 (u_P + u_N).n of the given field. */
int nbl_octree_flow_set_state(nbl_octree_flow* f, const double* u, const double* v, const double* w,
                              const double* p);
int nbl_octree_flow_get_state(nbl_octree_flow* f, double* u, double* v, double* w, double* p);
/* One step; dt <= 0 takes the CFL step. */
int nbl_octree_flow_step(nbl_octree_flow* f, double dt, nbl_octree_step_report* rep);
/* sum_f U_f A_f / V per leaf (caller order; 0 on non-fluid leaves) of the
 * current face velocities. */
int nbl_octree_flow_divergence(nbl_octree_flow* f, double* div);

/* No counterpart in the reference: measurement.  Average ms and algorithmic
 * bytes of one launch of an octree kernel (kernel 0: the SpMV, reading
 * coupling counts, masks, x and writing y once per leaf plus a column index
 * and a coefficient per coupling, or with coefficient classes
 * (nbl_octree_coef_classes) 8 bytes of classes and count per leaf, a column
 * index per coupling and the coefficients of couplings beyond a row's
 * sixth; kernel 1: the transport step, per leaf its
 * level, face counts, first-neighbour slots, velocities and phi in and out,
 * plus an index per further neighbour; kernel 2: the sensors, per leaf its
 * centre, level, mask, face counts, first-neighbour slots, u, v, w, p, the
 * carried state, the permutation, ten results and three flags, plus an index
 * per further neighbour), and the device time of construction (sort,
 * validation, face table, operator) and of the last solve, transport or
 * sensor evaluation. */
int nbl_octree_profile(nbl_octree* oc, int kernel, int reps, double* ms, double* bytes);
int nbl_octree_last_timing(nbl_octree* oc, double* setup_ms, double* solve_ms);

/* No counterpart in the reference: measurement.  In-situ timing of the finest
 * level's plain smoother sweeps inside this run's pressure solves: every
 * block of such a launch stores the device wall clock at its start and when
 * all its waves are done, and a launch lasts from its first start to its last
 * end (the first 2048 timed launches after enabling; launches a converged
 * solve skips are not counted).  This is the kernel's execution span; a
 * profiler's dispatch timestamps add the command processor's dispatch and
 * completion handling.  enable: 1 on (resets the totals), 0 off (keeps them),
 * -1 leave as is.  Reading the totals synchronises the device.
 * Reports the average ms per launch, the number of launches timed and the
 * sweep's algorithmic bytes (x, b, x' and the 2-byte code per cell). */
int nbl_run_sweep_timing(nbl_run* run, int enable, double* avg_ms, long long* launches, double* bytes);

/* ---- serving: device-resident request loop -------------------------------
 * No counterpart in the reference: its backend (backend/app/jobs.py
 * LocalJobRunner) runs one nabla_solve subprocess per job.  A server owns many
 * runs, each resident in HBM with its own stream, and `slots` worker threads
 * that execute step requests (nbl_run_step with the CFL dt) in arrival order,
 * letting later requests for idle runs overtake a request whose run is busy,
 * so different runs' steps overlap on the device.  Results do not depend on
 * what else is in flight.  Admission builds the run, measures its device
 * footprint and refuses it (NBL_ERR_CAPACITY) if the total would exceed
 * `hbm_budget` bytes; admission allocates and should happen outside latency-
 * critical windows.  Runs returned by nbl_server_run belong to the server
 * (never nbl_run_destroy them) and may be used with nbl_run_* only while they
 * have no request in flight.  A ticket is collected exactly once by
 * nbl_server_wait, which reports the step and its latency (submit to
 * completion, ms). */
int nbl_server_create(nbl_ctx* ctx, int slots, uint64_t hbm_budget, nbl_server** out);
int nbl_server_destroy(nbl_server* srv);   /* finishes queued requests first */
int nbl_server_admit(nbl_server* srv, const nbl_grid_desc* grid, const nbl_flow_cfg* cfg,
                     const nbl_solver_opts* opts, int* run_id);
int nbl_server_run(nbl_server* srv, int run_id, nbl_run** run);
/* An octree flow run (nbl_octree_flow_*): the leaves as nbl_octree_create
 * takes them and the flow configuration; admission also builds the octree's
 * multigrid hierarchy, so the measured footprint is the run's whole device
 * memory.  Its steps are served like the uniform runs' (dt from the CFL
 * rule); nbl_server_wait reports them in nbl_step_report's fields (cd, cl and
 * the momentum/continuity residuals 0, mass_error = inflow - outflow,
 * cell_count = leaves).  nbl_server_octree_flow returns the run's flow state
 * for set_state / get_state while it has no request in flight; it belongs to
 * the server (never nbl_octree_flow_destroy it).  nbl_server_run refuses
 * octree runs and nbl_server_octree_flow uniform ones. */
int nbl_server_admit_octree(nbl_server* srv, const uint64_t* keys, const uint8_t* levels,
                            const uint8_t* masks, size_t n, const double origin[3],
                            const double extent[3], int max_level, const nbl_octree_flow_cfg* cfg,
                            int* run_id);
int nbl_server_octree_flow(nbl_server* srv, int run_id, nbl_octree_flow** flow);
int nbl_server_submit(nbl_server* srv, int run_id, uint64_t* ticket);
int nbl_server_wait(nbl_server* srv, uint64_t ticket, nbl_step_report* rep, double* latency_ms);
int nbl_server_footprint(nbl_server* srv, uint64_t* used, uint64_t* budget);

/* Batched pressure solves: n independent solves (each solver as created by
 * nbl_poisson_create; host vectors), at most `concurrency` in flight at once
 * on the solvers' own streams.  Same per-problem results as n calls of
 * nbl_poisson_solve.  No counterpart in the reference. */
int nbl_poisson_solve_batch(nbl_poisson* const* ps, const double* const* b, double* const* x, int n,
                            int concurrency, nbl_poisson_report* reps);
/* Batched pressure solves (no counterpart in the reference, whose jobs each
 * run MgPoisson::solve on their own, backend/app/jobs.py:50-140): `problems`
 * independent systems on one grid (the same operator, different right-hand
 * sides), solved together by MgPoisson::solve's flexible CG (kind 0,
 * multigrid.cpp:267-355): every level kernel is one launch for all problems
 * (their vectors one after another in HBM, one copy of the operator codes
 * per problem, one coarsest-level inverse), while each problem keeps its own
 * scalars, stopping test and report, so a problem's arithmetic is that of a
 * batch of one.  The batch runs the one-kernel-per-sweep V-cycle with a
 * direct coarsest solve, FP64 or FP32 levels (opts precision NBL_PREC_FP64
 * or NBL_PREC_MIXED), levels in plain launch order or bands of rows; its
 * reductions are finalized per problem (never folded inline).  Other
 * options are refused with NBL_ERR_INVALID, as is a batch handle passed to
 * a single-problem nbl_poisson_* entry point.  Destroy with
 * nbl_poisson_destroy. */
int nbl_poisson_batch_create(nbl_ctx* ctx, const nbl_grid_desc* grid, int problems, double tol,
                             int max_iters, const nbl_solver_opts* opts, nbl_poisson** out);
/* Problems of a batch (0: a single-problem solver). */
int nbl_poisson_batch_size(nbl_poisson* ps, int* problems);
/* Host arrays: b[q], x[q] of the grid's cell count for each problem q;
 * reps[q] (may be NULL) its report. */
int nbl_poisson_batch_solve(nbl_poisson* ps, const double* const* b, double* const* x,
                            nbl_poisson_report* reps);
/* Device arrays holding the problems one after another (problems x cells,
 * 16-byte aligned). */
int nbl_poisson_batch_solve_device(nbl_poisson* ps, const double* d_b, double* d_x,
                                   nbl_poisson_report* reps);
/* The same for octree Jacobi-PCG solves (as nbl_octree_poisson_solve each). */
int nbl_octree_solve_batch(nbl_octree* const* os, const double* const* b, double* const* x, int n,
                           int concurrency, double tol, int max_iters, nbl_poisson_report* reps);
/* The same batch with nbl_octree_mg_solve (the hierarchy of each octree is
 * built on its first multigrid solve). */
int nbl_octree_mg_solve_batch(nbl_octree* const* os, const double* const* b, double* const* x, int n,
                              int concurrency, double tol, int max_iters, nbl_poisson_report* reps);

#ifdef __cplusplus
}
#endif
#endif /* NBL_API_H */
