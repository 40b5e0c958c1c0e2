// solver.hpp — host-side objects behind the C-ABI: device context, the
// uniform pressure solvers (MG-PCG and Jacobi-CG) and the MAC run.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dev.hpp"
#include "nbl_api.h"

struct nbl_ctx {
  int device = 0;
  std::atomic<uint64_t> bytes{0};
};

namespace nbl {

// Device allocation owned by an object and accounted to its context.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(nbl_ctx* ctx, size_t bytes);
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { swap(o); }
  DevBuf& operator=(DevBuf&& o) noexcept {
    swap(o);
    return *this;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p_);
  }
  size_t bytes() const { return n_; }

 private:
  void swap(DevBuf& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    std::swap(ctx_, o.ctx_);
  }
  void* p_ = nullptr;
  size_t n_ = 0;
  nbl_ctx* ctx_ = nullptr;
};

// Host copy of a uniform grid (reference UniformGrid).
struct HostGrid {
  int nx = 1, ny = 1, nz = 1;
  double org[3] = {0, 0, 0};
  double dx = 1, dy = 1, dz = 1;
  nbl_face_bc face[6]{};
  std::vector<uint8_t> solid;
  size_t cells() const { return size_t(nx) * ny * nz; }
  static HostGrid from(const nbl_grid_desc* d);
};

// One level of the agglomeration hierarchy, host side.
struct HostLevel {
  HostGrid g;
  std::vector<uint16_t> code;
  long long nf = 0;
  int f[3] = {1, 1, 1};
};

// Builds the levels as MgPoisson does (multigrid.cpp:210-266).
std::vector<HostLevel> buildLevels(const HostGrid& g, bool multilevel);
bool hasOutlet(const HostGrid& g);

// Uniform pressure solver on the device.  kind 0 = flexible CG with one
// V(pre,post) cycle per iteration (MgPoisson), kind 1 = Jacobi CG (CgPoisson).
class PoissonDev {
 public:
  // problems > 0: a batch of that many independent problems on the same grid,
  // solved together (one launch per level kernel for all of them; each
  // problem has its own scalars, convergence test and report).  Vectors are
  // then the problems' vectors one after another.
  PoissonDev(nbl_ctx* ctx, const HostGrid& g, int kind, double tol, int maxIters,
             const nbl_solver_opts& opts, int problems = 0);
  ~PoissonDev();
  // Solve with device buffers on `stream` (b and x may be any device arrays of
  // length n).  Returns the reference's Report.
  nbl_poisson_report solve(const double* d_b, double* d_x, hipStream_t stream);
  int problems() const { return nb_; }
  bool batched() const { return batched_; }
  // Reports of the last solve, one per problem.
  const std::vector<nbl_poisson_report>& reports() const { return reps_; }
  // One application of the preconditioner (device vectors).
  void precondition(const double* d_r, double* d_z, hipStream_t stream);
  const std::vector<HostLevel>& levels() const { return hl_; }
  const MgLevel& level(int l) const { return dl_[l]; }
  int nlevels() const { return int(dl_.size()); }
  bool directCoarse() const { return direct_; }
  bool singular() const { return singular_; }
  int kind() const { return kind_; }
  double lastMs() const { return lastMs_; }
  int lastVcycles() const { return lastV_; }
  RedBuf red() const { return RedBuf{red_.as<double>(), redStride_}; }
  int blocks(int l) const { return blocks_[l]; }
  bool projects() const { return project_; }
  // In-situ timing of the finest level's plain sweeps inside solves (device
  // wall-clock stamps per block; on: resets the totals).  sweepTiming folds
  // the stamps recorded so far (it synchronises the device).
  void setSweepTiming(bool on);
  void sweepTiming(double* ms, long long* n, double* bytes);
  // Average ms per launch and algorithmic bytes of one solve kernel.
  void profile(int kernel, int reps, hipStream_t s, double* ms, double* bytes);

 private:
  template <class T>
  void vcycle(int l, const T* b, MeanRef mb, const T** xout, MeanRef* mxout, hipStream_t s);
  // z = Minv r.  rOld: fold the flexible-CG dots z.r, z.rOld into the output
  // pass.  fused: the level-0 first sweep (and FP32 right-hand side) were
  // written by the update, and the last level-0 sweep writes z and the dots.
  // b0Ready: the FP32 level-0 right-hand side is already written (mixed)
  // fused: the level-0 update wrote the first sweep and the last sweep writes
  // z and the flexible-CG dots; x1Ready: the first sweep alone was written
  void minv(const double* r, hipStream_t s, const double* rOld, bool fused, bool b0Ready = false,
            bool x1Ready = false);
  bool buildDirectCoarse();
  // row0: the first partial row the finalize folds
  void fin(int op, int K, int nparts, int slot, long long nf, bool predicated, hipStream_t s, int earlyExit = 0,
           double convTol = 0.0, int row0 = 0);
  // Pending means of the null-space projections.  The mean in slot i of a
  // producer with `blocks` blocks is folded by its consumers from the slot's
  // own partial row when `inl` (few enough partials), else finalized into
  // SolveCtl by k_fin.
  bool inlineMean(int blocks) const { return !batched_ && blocks <= kInlineFoldMax; }
  MeanRef meanRef(int slot, bool inl, int blocks, long long nf) const;
  RedBuf meanRed(int slot, bool inl) const;
  MeanRef zMean() const;  // pending mean of the preconditioner output z
  void meanFin(int slot, bool inl, int blocks, long long nf, bool predicated, hipStream_t s);
  void enqueueIteration(hipStream_t s, int k);
  // The V(2,2) cycle in two kernels per level (vcycle.hip).  first: the
  // cycle applied to b0 (the first of a solve, precondition); otherwise
  // iteration k's, whose level-0 right-hand side is r_new of the update
  // (formed by the down kernel from r, A p and alpha; on singular systems read
  // from rnRaw_ with its pending mean).  b0: the level-0 right-hand side as
  // the up kernel reads it.
  template <class T>
  void tiledCycle(hipStream_t s, int k, bool first, const T* b0, const double* r, double* rn, const double* p,
                  InFold alpha);
  void enqueueIterationTiled(hipStream_t s, int k);
  double* slotRow(int slot) const {
    return red_.as<double>() + size_t(kMaxRedVals + slot) * size_t(redStride_);
  }
  MeanRef zMeanTiled() const;
  // The reference's loop structure, one kernel per loop (opts.loops).
  void removeMeanLoops(int l, double* x, hipStream_t s);
  void smoothLoops(int l, const double* b, double* x, int sweeps, hipStream_t s);
  void vcycleLoops(int l, const double* b, double* x, hipStream_t s);
  void minvLoops(const double* r, hipStream_t s);
  void enqueueIterationLoops(hipStream_t s);
  void readCtl(hipStream_t s);
  // Flexible-CG iterations k, k + 1 (k odd) from the captured graph: their
  // launches depend on k only through k & 1, so one capture of iterations 1
  // and 2 replays every later pair.  Captured on first use; false: not in
  // use (iteration_graph other than 2, the reference loop structure, in-situ
  // sweep timing, or the capture failed).
  bool iterationPair(hipStream_t s);
  hipGraphExec_t iterGraph_ = nullptr;
  bool graphFailed_ = false;
  int graphV_ = 0;  // V-cycles one replay applies
  long long* timedSweep(int l, bool plain);

  nbl_ctx* ctx_;
  int kind_;
  double tol_;
  int maxIt_;
  int nb_ = 1;            // problems (1 unless batched)
  bool batched_ = false;  // created as a batch: every reduction through k_fin, per problem
  std::vector<nbl_poisson_report> reps_;
  nbl_solver_opts opts_;
  bool singular_ = false;
  bool project_ = false;  // singular with fluid cells: null-space projections apply
  bool mixed_ = false;
  bool bf16_ = false;   // mixed with the finest level's pre-smoothed iterate stored as BF16
  DevBuf xaB_;
  std::vector<HostLevel> hl_;
  std::vector<MgLevel> dl_;
  std::vector<int> blocks_;   // geo3 block count per level (= partials per reduction)
  std::vector<int> rblocks_;  // blocks of the restriction into level l (l >= 1)
  std::vector<DevBuf> codes_;
  // per level: b (levels >= 1), two x buffers, coarse scratch
  std::vector<DevBuf> lb_, lx0_, lx1_;
  DevBuf coarseScr_;
  static constexpr long long kDirectCoarseMax = 4096;  // cells of a coarsest level solved directly
  bool direct_ = false;
  DevBuf coarseM_, coarseMf_;  // coarsest-level inverse (FP64, FP32)
  DevBuf b0f_;  // FP32 level-0 right-hand side (mixed)
  // flexible-CG vectors (FP64)
  DevBuf x_, rhs_, r_, rn_, z_, p_, p2_, Ap_;
  // level-0 fusions of the current V-cycle (set by minv)
  bool l0Fused_ = false;
  bool l0First_ = false;  // the level-0 first sweep from zero is already in lx0_[0]
  bool l0FoldLast_ = false;  // the last level-0 sweep folds its fluid sum (z's pending mean)
  // Singular systems (one kernel per sweep, a coarser level, two or more
  // post-sweeps): the fused last level-0 sweep writes z with its mean pending
  // (partial rows kZDotsRow.., z.rn, z.r, fluid sum; folded inline or by
  // FIN_BETA_FLEX_MEAN) and the next A p removes it.
  bool zDotsMean() const {
    return kind_ == 0 && project_ && !tiled_ && !loops_ && nlevels() > 1 && opts_.pre_smooth >= 1 &&
           opts_.post_smooth >= 2 && dl_[0].nf > 0;
  }
  static constexpr int kZDotsRow = 4;
  const double* l0Rn_ = nullptr;
  const double* l0R_ = nullptr;
  // Finalizes folded into their consumers for the current solve (flexible CG
  // without projections, few enough level-0 blocks); conv_: the convergence
  // test the V-cycle's first level-0 sweep performs.
  static constexpr int kInlineFoldMax = 2048;
  // the two-kernels-per-level cycle folds every reduction inline (every block
  // reads all partials); it is used up to this many level-0 workgroups.
  // Measured on periodic boxes (mixed form, ms per step): 100^3 (2275 tiles)
  // 12.2 -> 9.9 and 116^3 (3480) 13.6 -> 12.8 against one kernel per sweep,
  // while 146^3 (7030 tiles) went from 24.4 to 29.2
  static constexpr int kTiledFoldMax = 4096;
  static constexpr int kSlotRows = 64;  // partial rows of the pending-mean slots, after the kMaxRedVals rows
  bool inline_ = false;
  InFold conv_{};
  bool loops_ = false;
  bool tiled_ = false;       // the two-kernels-per-level cycle is in use
  std::vector<int> tblocks_;  // workgroups of the tiled kernels per level
  DevBuf rnRaw_;             // singular tiled path: r_new before its projection
  std::vector<DevBuf> ldiag_;  // loops: per-level diagonal (CgPoisson::buildDiagonal)
  DevBuf ctl_, red_;
  int redStride_ = 1;
  SolveCtl* hctl_ = nullptr;  // pinned mirror
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  double lastMs_ = 0;
  int lastV_ = 0;
  int predicted_ = 1;
  int vcount_ = 0;
  static constexpr int kStampLaunches = 2048;  // timed launches held before a fold
  bool timeSweeps_ = false;
  DevBuf stamps_;  // per timed launch and block: start, end (device wall clock)
  int stampCap_ = 0, stampUsed_ = 0, wallKHz_ = 100000;
  double sweepMs_ = 0.0;
  long long sweepN_ = 0;
};

// A device-resident MAC Navier-Stokes run (NSSolver).
class RunDev {
 public:
  RunDev(nbl_ctx* ctx, const HostGrid& g, const nbl_flow_cfg& cfg, const nbl_solver_opts& opts);
  ~RunDev();
  void initialize();
  void setState(const double* u, size_t nu, const double* v, size_t nv, const double* w, size_t nw,
                const double* p, size_t np, int step, double t, double dtLast);
  void getState(double* u, double* v, double* w, double* p, int* step, double* t, double* dtLast);
  double computeDt();
  nbl_step_report step(double dt);
  void rhs(double* ru, double* rv, double* rw);
  void cellVelocity(int comp, double* out);
  double* u() const { return cur_[0]; }
  double* v() const { return cur_[1]; }
  double* w() const { return cur_[2]; }
  double* p() const { return p_.as<double>(); }
  hipStream_t stream() const { return stream_; }
  void lastSolves(double* ms, int* iters, int* vcycles) const {
    for (int q = 0; q < 3; ++q) {
      if (ms) ms[q] = solveMs_[q];
      if (iters) iters[q] = solveIt_[q];
      if (vcycles) vcycles[q] = solveV_[q];
    }
  }
  size_t sizes(int q) const { return sz_[q]; }
  PoissonDev& poisson() { return *poisson_; }
  // A host pressure solver in place of the device MG-PCG (NULL: device).
  void setHostSolver(nbl_host_poisson_fn fn, void* user) {
    hostFn_ = fn;
    hostUser_ = user;
  }

 private:
  void bc(bool outflow);
  void project(int stageIdx, nbl_poisson_report& rep);

  nbl_ctx* ctx_;
  HostGrid g_;
  nbl_flow_cfg cfg_;
  int rhsPlanes_ = 0;  // WENO-5 brick march length (0 automatic)
  NsArgs args_{};
  hipStream_t stream_ = nullptr;
  size_t sz_[4];
  DevBuf solid_, solidList_;
  int solidCounts_[3] = {0, 0, 0};
  DevBuf bufA_[3], bufB_[3], buf0_[3], p_, b_, phi_;
  double* cur_[3];
  double* alt_[3];
  DevBuf sc_, rep_, red_, gridDev_;
  DevBuf padBuf_[3];
  PadBufs pad_{};
  int redStride_ = 1;
  StepCtl* hsc_ = nullptr;
  DevReport* hrep_ = nullptr;
  std::unique_ptr<PoissonDev> poisson_;
  nbl_host_poisson_fn hostFn_ = nullptr;
  void* hostUser_ = nullptr;
  std::vector<double> hostB_, hostX_;
  double dtLast_ = 0.0;
  double solveMs_[3] = {0, 0, 0};
  int solveIt_[3] = {0, 0, 0}, solveV_[3] = {0, 0, 0};
};

}  // namespace nbl

struct nbl_run {
  nbl_ctx* ctx;
  std::unique_ptr<nbl::RunDev> impl;
};
struct nbl_poisson {
  nbl_ctx* ctx;
  std::unique_ptr<nbl::PoissonDev> impl;
  hipStream_t stream = nullptr;
  nbl::HostGrid grid;
};
