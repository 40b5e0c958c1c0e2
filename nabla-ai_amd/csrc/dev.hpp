// dev.hpp — POD descriptors shared by host code and gfx950 kernels, and the
// host-side launchers each kernel file exports.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace nbl {

// Uniform MAC grid as the kernels see it (reference grid.hpp:32-73).
struct DevGrid {
  int nx, ny, nz;
  double dx, dy, dz;
  double ox, oy, oz;
  int bc[6];              // NBL_BC_* per face 0=-x .. 5=+z
  double wall[6][3];      // wall velocity per face
  int profile[6];         // inlet profile per face
  double inletSpeed[6];
  int inletAxis[6];
  const uint8_t* solid;   // device, nx*ny*nz
  int hasSolid;
};

// One multigrid level.  code[c]: bits 2d..2d+1 = coupling kind across face d
// (0 none, 1 fluid neighbour, 2 outlet Dirichlet; d = 0..5 = -x,+x,-y,+y,-z,+z),
// bit 12 = fluid cell.  The diagonal is rebuilt from the code in the same
// accumulation order as the reference's diagOf (poisson.cpp:54-74).
struct MgLevel {
  int nx, ny, nz;
  long long n;
  double h2[3];           // 1/dx^2, 1/dy^2, 1/dz^2
  const uint16_t* code;
  int fx, fy, fz;         // factor from the next-finer level (coarse levels)
  long long nf;           // fluid cells
  int vec;                // cells per thread along i in the cell kernels (1, 2 or 4; divides nx)
  int slab;               // launch order of the cell kernels (blk3): 0 plain, 1 XCD slabs of planes, 2 XCD bands of rows
  int nb;                 // problems of a batch: level vectors and codes hold nb blocks of n cells, a
                          // launch's grid z runs over nb * nz planes (0 or 1: one problem)
};

constexpr uint16_t kFluidBit = 1u << 12;

__host__ __device__ inline int kindOf(uint16_t code, int dir) { return (code >> (2 * dir)) & 3; }

// Scalars of one flexible-CG / CG solve, resident on the device.  Kernels of
// an iteration read `done` first and do nothing once the solve has stopped,
// so the host may enqueue iterations ahead of the convergence test.
struct SolveCtl {
  double tol, bnorm, rz, pAp, alpha, beta, rr, resid;
  int it, done, converged, broke;
  int early, pad_;  // early: converged before the first iteration
  double rzq[2];    // r.z of flexible-CG iteration k in rzq[k & 1] (finalizes folded into kernels)
  double slot[64];  // reduction results / pending means
};

// A pending null-space mean (fluid sum / nf) that a kernel subtracts on read:
// either finalized into a SolveCtl slot by k_fin (`slot`), or left as the
// producer's `n` per-block partial sums, which every block of the consumer
// folds itself (`parts`, as InFold).  Neither: no mean.
struct MeanRef {
  const double* slot = nullptr;
  const double* parts = nullptr;
  int n = 0;
  double nf = 1.0;
  __host__ __device__ bool on() const { return slot != nullptr || parts != nullptr; }
  // The same slot in problem p's SolveCtl (batches: one SolveCtl per problem,
  // contiguous; batches use slots only).
  __host__ __device__ MeanRef at(int p) const;
};
constexpr int kCtlDoubles = int(sizeof(SolveCtl) / sizeof(double));
static_assert(sizeof(SolveCtl) % sizeof(double) == 0, "SolveCtl must be a whole number of doubles");
__host__ __device__ inline MeanRef MeanRef::at(int p) const {
  MeanRef r = *this;
  if (r.slot) r.slot += size_t(p) * size_t(kCtlDoubles);
  return r;
}
__host__ __device__ inline SolveCtl* ctlAt(SolveCtl* c, int p) { return c ? c + p : c; }
__host__ __device__ inline const SolveCtl* ctlAt(const SolveCtl* c, int p) { return c ? c + p : c; }
inline MeanRef slotMean(const double* slot) {
  MeanRef m;
  m.slot = slot;
  return m;
}

// A finalize folded into the kernel that consumes it, in place of a one-block
// k_fin launch (flexible CG on systems without projections): every block of
// the consumer folds the previous kernel's `n` per-block partials (rows of
// `stride`) in the same fixed order, so all blocks reach the same bits, and
// block 0 records the scalars in SolveCtl.  Scalars that a consumer reads
// while its own block 0 writes their successor alternate by iteration parity
// (SolveCtl::rzq).  n == 0: off, the scalars come from SolveCtl.
struct InFold {
  const double* parts = nullptr;
  int stride = 0, n = 0;
  int k = 0;  // flexible-CG iteration
};

// Launch shape of a cell kernel on an nx x ny x nz level: 256-thread blocks,
// bx threads along i (64, 32, 16 or 8) times 256/bx rows of j, one k per grid z.
struct Geo3 {
  dim3 grid, block;
  int blocks;
};
Geo3 geo3(int nx, int ny, int nz);
// `cells` consecutive cells along i per thread; nb problems of a batch stacked
// along the grid's z (plain order or bands of rows; blocks counts them all)
Geo3 geo3n(int nx, int ny, int nz, int cells, int slab = 0, int nb = 1);
// Automatic launch order of a level's cell kernels (blk3's slab argument): XCD
// slabs of planes (1) on levels below 8M cells with at least 8 planes, bands
// of rows (2) from 8M cells, plain order (0) otherwise.  On the 192 x 64 x 64
// channel slabs make the step 3% faster than plain order; on 256^3, 512 x 256
// x 128 and 320 x 320 x 160 bands run A p 3-4% faster than plain order (78-79
// vs 81 us) and the FP64 sweep equally fast, while slabs of planes are slower
// than both (sweep 99-102 vs 93-95 us).
int slabFor(int nx, int ny, int nz);
Geo3 levelGeo(const MgLevel& L);                // the launch of a level's cell kernels (L.vec)
// Launch shape of the fused residual + restriction from L to Lc: one thread per
// fine cell, blocks covering whole coarse cells (its block count is the
// number of partials the fold writes).
Geo3 restrictGeo(const MgLevel& L, const MgLevel& Lc);

// Scalar update applied by the one-block finalize kernel after a reduction.
enum FinOp : int {
  FIN_MEAN = 0,   // slot[slot] = sum / nf
  FIN_BNORM,      // |b| -> bnorm, reset counters, optional early exit
  FIN_RZ,         // rz = r.z
  FIN_ALPHA,      // p.Ap -> alpha (or stop on a vanishing p.Ap)
  FIN_CONV,       // |r|^2 -> convergence test
  FIN_BETA_FLEX,  // (z.rn - z.r)/rz
  FIN_BETA,       // (r.z)/rz
  FIN_RESID,      // |rhs - A x|^2 -> true relative residual
  FIN_BNORM_MEAN, // (|b|^2, sum) -> bnorm + counters reset, slot[slot] = sum / nf
  FIN_CONV_MEAN,  // (|r|^2, sum) -> convergence test, slot[slot] = sum / nf
  FIN_BNORM_R0,   // (|b|^2, |r0|^2) -> bnorm, reset counters, optional early exit on |r0|
  FIN_BETA_FLEX_MEAN,  // (z.rn, z.r, z's fluid sum) -> FIN_BETA_FLEX, slot[slot] = sum / nf
};
struct FinArgs {
  int op;
  int K;          // values folded (1 to 3)
  int nparts;     // partials written by the reducing kernel (its block count)
  int slot;
  long long nf;
  double convTol;
  int earlyExit;
  int predicated; // skip when the solve has stopped
};

// Per-run time-step scalars (NSSolver step state).
struct StepCtl {
  double dt;       // dt of the current step (CFL or fixed)
  double t;
  int step;
  int pad_;
  double cfl_max;  // max inverse CFL time scale found by the last CFL pass
};

// Step report fields computed on the device (mirrors nbl_step_report).
struct DevReport {
  int step;
  double t, dt, cfl, mom, divl2, divmax, mass;
  int piters;
  double pres, cd, cl;
  unsigned long long cells;
  int accepted;
};

// ---- launchers (ns.hip) ----------------------------------------------------
struct NsArgs {
  DevGrid g;
  double nu;
  double bodyForce[3];
  int convection;
  double cfl, velocityScale, density, refLength, refArea, tolContinuity, tolMass;
};

// Padded copies of u, v, w used by the RHS stencils: kPadHalo ghost layers on
// every side.  Element (i,j,k) of component c sits at
// p[c][origin[c] + i + sx[c]*j + sxy[c]*k]; n[c] elements are allocated.
constexpr int kPadHalo = 3;
struct PadBufs {
  double* p[3];
  long long origin[3];
  int sx[3], sxy[3];
  long long n[3];
};
PadBufs padLayout(const DevGrid& g);  // sizes and strides; the caller sets p[]

// Brick tiling of the WENO-5 right-hand side (ns.hip): per component the box
// of faces that can be active (origin, extents), its 64 x 4 x planes bricks,
// and the faces outside the boxes (right-hand side zero) per component.
constexpr int kBrickX = 64, kBrickY = 4;
struct BrickComp {
  int ox, oy, oz, ex, ey, ez;
  int tx, ty, tz, blocks;
};
struct BrickArgs {
  BrickComp c[3];
  int planes, tileBlocks;
  long long edge[3];
};
BrickArgs brickArgs(const DevGrid& g, int planes);

void launch_bc_planes(const NsArgs& a, double* u, double* v, double* w, bool outflow, hipStream_t s);
void launch_bc_solid(const NsArgs& a, double* u, double* v, double* w, const int* list,
                     const int counts[3], hipStream_t s);
// Pads u, v, w (gd: device copy of a.g) and writes the stage update: stage 0
// writes R itself, stages 1-3 the SSP-RK3 combinations with u0.
void launch_rhs_rk(const NsArgs& a, const DevGrid* gd, const PadBufs& pb, const double* u0, const double* v0,
                   const double* w0, const double* u, const double* v, const double* w, double* uo, double* vo,
                   double* wo, int stage, const StepCtl* sc, hipStream_t s, int planes = 0);
void launch_divergence(const NsArgs& a, const double* u, const double* v, const double* w,
                       double* b, hipStream_t s);
void launch_correct(const NsArgs& a, double* u, double* v, double* w, const double* phi,
                    double* p, const StepCtl* sc, hipStream_t s);
// Reducing NS launches write partials and run their own finalize kernel.
void launch_cfl(const NsArgs& a, const double* u, const double* v, const double* w, StepCtl* sc,
                double fixed_dt, RedBuf rb, hipStream_t s);
int ns_partials(const DevGrid& g);  // partial slots the NS reductions need
void launch_diagnostics(const NsArgs& a, const double* u, const double* v, const double* w,
                        const double* u0, const double* v0, const double* w0, const double* p,
                        StepCtl* sc, DevReport* out, RedBuf rb, hipStream_t s);
void launch_cell_velocity(const NsArgs& a, const double* u, const double* v, const double* w,
                          int comp, double* out, hipStream_t s);
void launch_fill(double* a, long long n, double v, hipStream_t s);

// ---- launchers (mg.hip) ----------------------------------------------------
// Multigrid pieces are templated on the storage/arithmetic type T (double for
// the plain path, float for the mixed V-cycle).
// nb > 1: one block per problem p of a batch, folding the producer's
// partials p * a.nparts .. (p + 1) * a.nparts - 1 into problem p's SolveCtl
void launch_fin(RedBuf rb, const FinArgs& a, SolveCtl* ctl, hipStream_t s, int nb = 1);
// All solve scalars to zero, tolerance set (a kernel, ordered on `s`).
void launch_ctl_init(SolveCtl* ctl, double tol, hipStream_t s, int nb = 1);
// Step scalars (step index, time, last dt) set by a kernel on `s`.
void launch_set_step(StepCtl* sc, int step, double t, double dt, hipStream_t s);
// stamps: optional per-block device wall-clock stamps (start, and end once
// the block's stores are visible), 2 per block.  conv: the flexible CG's
// convergence test on the partials of |r_new|^2 folded in (the sweep does
// nothing when it stops the solve).
template <class T>
void launch_mg_sweep(const MgLevel& L, const T* b, MeanRef mb, const T* x, MeanRef mx,
                     bool zeroGuess, const T* ec, MeanRef mec, const MgLevel* Lc, T* xo, bool fold,
                     RedBuf rb, SolveCtl* ctl, hipStream_t s, long long* stamps = nullptr, InFold conv = {});
// x1c (optional): the coarse level's first sweep from a zero guess, written
// beside the restricted residual (pointwise: no projection may come between,
// so only where the restricted residual is not projected).
template <class T>
void launch_mg_resid_restrict(const MgLevel& L, const MgLevel& Lc, const T* b, MeanRef mb,
                              const T* x, MeanRef mx, T* rc, bool fold, RedBuf rb,
                              SolveCtl* ctl, hipStream_t s, T* x1c = nullptr, InFold conv = {});
template <class T>
void launch_mg_coarse(const MgLevel& L, bool singular, const T* b, const double* mb, T* x,
                      T* scratch, const SolveCtl* ctl, hipStream_t s);
// Direct coarse solve: one Gauss-Jordan step on the n x 2n augmented matrix
// (sets *fail when a pivot is not above `tiny`), the projected inverse from the
// eliminated matrix (scratch: n + 1 doubles), and x = Minv (b - mb) per solve
// (FP32 levels use the FP32 copy Mf).
void launch_gj_step(double* G, int n, int k, double* colk, double* rowk, int* fail, double tiny, hipStream_t s);
void launch_inv_store(const double* G, int n, const uint16_t* code, long long nf, bool singular, double* scratch,
                      double* M, float* Mf, hipStream_t s);
template <class T>
void launch_mg_coarse_direct(const MgLevel& L, const double* M, const float* Mf, const T* b, MeanRef mb, T* x,
                             const SolveCtl* ctl, hipStream_t s);
template <class T>
void launch_mg_finish(const MgLevel& L, const T* x, MeanRef mx, double* z, bool fold,
                      RedBuf rb, const SolveCtl* ctl, hipStream_t s);
template <class T>
void launch_to_level0(const MgLevel& L, const double* r, T* b, const SolveCtl* ctl, hipStream_t s);

// ---- the tiled V-cycle (vcycle.hip): two kernels per level ------------------
// Down kernel sources of the level's right-hand side: DS_B a level vector b
// (pending mean mb); DS_UPD level 0 of a flexible-CG iteration, r_new = r -
// alpha A p formed on the fly (x += alpha p and r_new written on the tile);
// DS_RN level 0 of a singular iteration, r_new read with its pending mean
// and written projected.
enum DownSrc : int { DS_B = 0, DS_UPD = 1, DS_RN = 2 };
template <class T>
struct DownArgs {
  MgLevel L, Lc;
  const T* b = nullptr;
  MeanRef mb{};
  const double* r = nullptr;
  const double* Ap = nullptr;
  double* x = nullptr;
  const double* p = nullptr;
  InFold alpha{};             // DS_UPD: alpha = rz / p.Ap folded from the partials of p.Ap
  const double* rnRaw = nullptr;
  MeanRef mrn{};              // DS_RN
  double* rnOut = nullptr;    // DS_UPD / DS_RN: r_new (projected for DS_RN)
  T* b0Out = nullptr;         // DS_UPD / DS_RN with T = float: the level-0 right-hand side float(r_new)
  T* xa = nullptr;            // x after the two pre-sweeps
  __hip_bfloat16* xaB = nullptr;  // the same stored as BF16 instead (finest level of the BF16 form)
  T* bc = nullptr;            // restricted residual
  double* rowNorm = nullptr;  // per-block |r_new|^2 (DS_UPD / DS_RN)
  double* rowMean = nullptr;  // per-block fluid sum of bc (singular systems)
  SolveCtl* ctl = nullptr;
  InFold conv{};              // the flexible CG's convergence test, folded at entry
};
// Up kernel destinations: UD_X the level's x; UD_Z level 0, z (and per-block
// fluid sums of z in rb when rb.partials is set); UD_DOTS level 0, z and the
// partials of z.r_new, z.r (rows 0, 1 of rb).
enum UpDst : int { UD_X = 0, UD_Z = 1, UD_DOTS = 2 };
template <class T>
struct UpArgs {
  MgLevel L, Lc;
  const T* b = nullptr;
  MeanRef mb{};
  const T* xa = nullptr;      // x after the pre-sweeps
  const __hip_bfloat16* xaB = nullptr;  // or the same stored as BF16
  const T* e = nullptr;       // the coarse level's correction
  T* xb = nullptr;            // UD_X
  double* z = nullptr;        // UD_Z / UD_DOTS
  const double* rn = nullptr;
  const double* r = nullptr;
  RedBuf rb{nullptr, 0};
  SolveCtl* ctl = nullptr;
  int countIt = 0;            // block 0 counts the finished flexible-CG iteration
};
template <class T>
void launch_down(int src, const DownArgs<T>& a, hipStream_t s);
template <class T>
void launch_up(int dst, const UpArgs<T>& a, hipStream_t s);
int tileBlocks(const MgLevel& L);           // workgroups (= partials) of a level's tiled kernels
bool tileFits(const MgLevel& L, const MgLevel& Lc);

// The reference's multigrid / flexible-CG loops one kernel per loop (the
// unfused FP64 baseline).  Reductions write one partial per block of
// loopBlocks(n) blocks.
constexpr int kLoopBlocks = 2048;
int loopBlocks(long long n);
void launch_loop_apply(const MgLevel& L, const double* x, double* out, const SolveCtl* ctl, hipStream_t s);
void launch_loop_jacobi(const MgLevel& L, const double* b, const double* Ax, const double* diag, double* x,
                        const SolveCtl* ctl, hipStream_t s);
void launch_loop_msum(const MgLevel& L, double* x, RedBuf rb, const SolveCtl* ctl, hipStream_t s);
void launch_loop_msub(const MgLevel& L, double* x, const double* mean, const SolveCtl* ctl, hipStream_t s);
void launch_loop_resid(const MgLevel& L, const double* b, double* r, const SolveCtl* ctl, hipStream_t s);
void launch_loop_restrict(const MgLevel& L, const MgLevel& Lc, const double* rf, double* rc, const SolveCtl* ctl,
                          hipStream_t s);
void launch_loop_prolong(const MgLevel& L, const MgLevel& Lc, const double* ec, double* xf, const SolveCtl* ctl,
                         hipStream_t s);
void launch_loop_dot(long long n, const double* a, const double* b, const double* b2, RedBuf rb,
                     const SolveCtl* ctl, hipStream_t s);
void launch_loop_update(long long n, double* x, const double* p, const double* r, const double* Ap, double* rn,
                        const SolveCtl* ctl, hipStream_t s);
void launch_loop_pupdate(long long n, double* p, const double* z, const SolveCtl* ctl, hipStream_t s);
void launch_loop_copy(long long n, double* dst, const double* src, const SolveCtl* ctl, hipStream_t s);

void launch_apply(const MgLevel& L, const double* x, double* out, hipStream_t s);
void launch_diag(const MgLevel& L, double* out, hipStream_t s);
void launch_restrict_plain(const MgLevel& L, const MgLevel& Lc, const double* rf, double* rc,
                           hipStream_t s);
void launch_prolong_plain(const MgLevel& L, const MgLevel& Lc, const double* ec, double* xf,
                          hipStream_t s);

// Flexible CG (MgPoisson::solve) and Jacobi CG (CgPoisson::solve) pieces.
// Each reducing launcher writes one partial per block of levelGeo(level 0).
// rhs = masked b, x = 0, partials of rhs's fluid sum; with r (no projection):
// r = rhs and partials of |rhs|^2 instead.  b0 / x1f / x1d (optional, need r
// or init2): the first V-cycle's FP32 right-hand side and its level-0 first
// sweep from zero.
void launch_fcg_init(const MgLevel& L, bool maskSolid, const double* b, double* rhs, double* x, RedBuf rb,
                     hipStream_t s, double* r = nullptr, float* b0 = nullptr, float* x1f = nullptr,
                     double* x1d = nullptr);
void launch_fcg_init2(const MgLevel& L, bool singular, double* rhs, double* r, const SolveCtl* ctl, RedBuf rb,
                      hipStream_t s, float* b0 = nullptr, float* x1f = nullptr, double* x1d = nullptr);
void launch_fcg_start(const MgLevel& L, double* z, MeanRef mz, const double* r, double* p,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s);
void launch_fcg_apply(const MgLevel& L, const double* p, double* Ap, const SolveCtl* ctl, RedBuf rb,
                      hipStream_t s);
void launch_fcg_update(const MgLevel& L, bool singular, double* x, const double* p, const double* r,
                       const double* Ap, double* rn, SolveCtl* ctl, RedBuf rb, hipStream_t s, InFold alpha = {});
// Non-singular systems: the update plus the next V-cycle's first level-0 sweep
// from zero (x1) and, for an FP32 cycle, its right-hand side b0 = float(rn).
// alpha: alpha = rz / p.Ap folded in from the partials of p.Ap.
template <class T>
void launch_fcg_update_first(const MgLevel& L, double* x, const double* p, const double* r, const double* Ap,
                             double* rn, T* x1, T* b0, SolveCtl* ctl, RedBuf rb, hipStream_t s, InFold alpha = {});
// p_new = z + beta p_old and A p_new, p_new.Ap in one pass.  beta: the
// flexible beta folded in from the partials of z.r_new and z.r.
// zMeanNf > 0: z's mean is pending and p_new = (z - mean) + beta p_old on
// fluid cells; beta's inline fold carries a third row, z's fluid sum, or
// (no inline fold) beta and the mean were finalized into ctl and zMeanSlot.
void launch_fcg_apply_pu(const MgLevel& L, const double* z, const double* pOld, double* pNew, double* Ap,
                         SolveCtl* ctl, RedBuf rb, hipStream_t s, InFold beta = {}, double zMeanNf = 0.0,
                         const double* zMeanSlot = nullptr);
// Last level-0 sweep of a V-cycle writing z and the partials of z.rn, z.r;
// countIt: the sweep also counts the finished iteration (block 0).
template <class T>
void launch_mg_sweep_dots(const MgLevel& L, const T* b, const T* x, double* z, const double* rn, const double* r,
                          RedBuf rb, SolveCtl* ctl, hipStream_t s, bool countIt = false, bool zSum = false);
// rn <- rn - mean (fluid) / 0 (solid), partials of |rn|^2.
// b0 (optional): also write float(r_new) there, the FP32 V-cycle's right-hand side
// x1f / x1d (optional, at most one): the level-0 first sweep from a zero guess on
// that right-hand side, in the cycle's type
void launch_fcg_norm(const MgLevel& L, double* rn, MeanRef mrn, const SolveCtl* ctl, RedBuf rb, hipStream_t s,
                     float* b0 = nullptr, float* x1f = nullptr, double* x1d = nullptr);
void launch_fcg_dots(const MgLevel& L, double* z, MeanRef mz, const double* rn, const double* r,
                     SolveCtl* ctl, RedBuf rb, hipStream_t s, bool countIt = false);
// k_finish fused into k_fcg_dots: z = V-cycle output x (minus its pending
// fluid mean mx); countIt: count the iteration (no finalize follows).
template <class T>
void launch_fcg_dots_x(const MgLevel& L, const T* x, MeanRef mx, double* z, const double* rn, const double* r,
                       SolveCtl* ctl, RedBuf rb, hipStream_t s, bool countIt = false);
void launch_fluid_sum(const MgLevel& L, const double* x, RedBuf rb, hipStream_t s);
void launch_fcg_final(const MgLevel& L, bool singular, const double* x, const double* rhs, double* xo,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s);
void launch_cg_z(const MgLevel& L, const double* r, double* z, bool first, double* p, const SolveCtl* ctl,
                 RedBuf rb, hipStream_t s);
void launch_cg_update(const MgLevel& L, double* x, const double* p, double* r, const double* Ap,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s);

}  // namespace nbl
