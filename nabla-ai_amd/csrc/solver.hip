// solver.hip — host orchestration of the device pressure solvers and the
// device-resident MAC run.  Kernel launches only; no per-element host work
// after set-up.
#include "solver.hpp"

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <stdexcept>

namespace nbl {

// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------
DevBuf::DevBuf(nbl_ctx* ctx, size_t bytes) : n_(bytes), ctx_(ctx) {
  if (bytes == 0) return;
  NBL_HIP(hipMalloc(&p_, bytes));
  if (ctx_) ctx_->bytes += bytes;
}
DevBuf::~DevBuf() {
  if (p_) {
    (void)hipFree(p_);
    if (ctx_) ctx_->bytes -= n_;
  }
}

// ---------------------------------------------------------------------------
// Grid and hierarchy (multigrid.cpp:210-266, poisson.cpp:18-52)
// ---------------------------------------------------------------------------
HostGrid HostGrid::from(const nbl_grid_desc* d) {
  if (!d) throw std::invalid_argument("grid descriptor is NULL");
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) throw std::invalid_argument("grid dimensions must be >= 1");
  if (!(d->dx > 0 && d->dy > 0 && d->dz > 0)) throw std::invalid_argument("grid spacings must be > 0");
  HostGrid g;
  g.nx = d->nx;
  g.ny = d->ny;
  g.nz = d->nz;
  for (int a = 0; a < 3; ++a) g.org[a] = d->origin[a];
  g.dx = d->dx;
  g.dy = d->dy;
  g.dz = d->dz;
  for (int f = 0; f < 6; ++f) {
    g.face[f] = d->face[f];
    if (g.face[f].type < NBL_BC_WALL || g.face[f].type > NBL_BC_PERIODIC)
      throw std::invalid_argument("unknown boundary type");
  }
  for (int a = 0; a < 3; ++a)
    if ((g.face[2 * a].type == NBL_BC_PERIODIC) != (g.face[2 * a + 1].type == NBL_BC_PERIODIC))
      throw std::invalid_argument("periodic boundaries must be paired");
  g.solid.assign(g.cells(), 0);
  if (d->solid) std::memcpy(g.solid.data(), d->solid, g.cells());
  return g;
}

bool hasOutlet(const HostGrid& g) {
  for (const auto& f : g.face)
    if (f.type == NBL_BC_OUTLET) return true;
  return false;
}

static void fillCodes(HostLevel& L) {
  const HostGrid& g = L.g;
  const int n[3] = {g.nx, g.ny, g.nz};
  L.code.assign(g.cells(), 0);
  L.nf = 0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t c = size_t(i) + size_t(g.nx) * (size_t(j) + size_t(g.ny) * k);
        if (g.solid[c]) continue;
        uint16_t code = kFluidBit;
        ++L.nf;
        for (int dir = 0; dir < 6; ++dir) {
          const int axis = dir / 2, side = (dir & 1) ? 1 : -1;
          if (axis == 2 && g.nz == 1) continue;
          int q[3] = {i, j, k};
          q[axis] += side;
          int kind;
          if (q[axis] >= 0 && q[axis] < n[axis]) {
            kind = 1;
          } else {
            const int t = g.face[2 * axis + (side > 0 ? 1 : 0)].type;
            if (t == NBL_BC_PERIODIC) {
              q[axis] = (q[axis] + n[axis]) % n[axis];
              kind = 1;
            } else {
              kind = t == NBL_BC_OUTLET ? 2 : 0;
            }
          }
          if (kind == 1 && g.solid[size_t(q[0]) + size_t(g.nx) * (size_t(q[1]) + size_t(g.ny) * q[2])]) kind = 0;
          code |= uint16_t(kind << (2 * dir));
        }
        L.code[c] = code;
      }
}

std::vector<HostLevel> buildLevels(const HostGrid& g, bool multilevel) {
  std::vector<HostLevel> lv;
  HostLevel L0;
  L0.g = g;
  fillCodes(L0);
  lv.push_back(std::move(L0));
  while (multilevel && lv.size() < 8) {
    const HostGrid& f = lv.back().g;
    const int fx = f.nx >= 4 ? 2 : 1, fy = f.ny >= 4 ? 2 : 1, fz = f.nz >= 4 ? 2 : 1;
    if (fx == 1 && fy == 1 && fz == 1) break;
    HostLevel C;
    C.f[0] = fx;
    C.f[1] = fy;
    C.f[2] = fz;
    HostGrid& cg = C.g;
    cg.nx = (f.nx + fx - 1) / fx;
    cg.ny = (f.ny + fy - 1) / fy;
    cg.nz = (f.nz + fz - 1) / fz;
    cg.dx = f.dx * double(f.nx) / double(cg.nx);
    cg.dy = f.dy * double(f.ny) / double(cg.ny);
    cg.dz = f.dz * double(f.nz) / double(cg.nz);
    for (int a = 0; a < 3; ++a) cg.org[a] = f.org[a];
    for (int q = 0; q < 6; ++q) cg.face[q] = f.face[q];
    std::vector<uint8_t> any(cg.cells(), 0);
    const HostLevel& F = lv.back();
    for (int k = 0; k < f.nz; ++k)
      for (int j = 0; j < f.ny; ++j)
        for (int i = 0; i < f.nx; ++i) {
          const size_t c = size_t(i) + size_t(f.nx) * (size_t(j) + size_t(f.ny) * k);
          if (F.code[c] & kFluidBit)
            any[size_t(i / fx) + size_t(cg.nx) * (size_t(j / fy) + size_t(cg.ny) * (k / fz))] = 1;
        }
    cg.solid.assign(cg.cells(), 0);
    for (size_t q = 0; q < cg.solid.size(); ++q) cg.solid[q] = any[q] ? 0 : 1;
    fillCodes(C);
    lv.push_back(std::move(C));
    if (lv.back().g.cells() <= 2048) break;
  }
  return lv;
}

// ---------------------------------------------------------------------------
// Pressure solver
// ---------------------------------------------------------------------------
namespace {
constexpr int kSlotMeanZ = 3;  // pending mean of the preconditioner output z
constexpr int kSlotMeanRhs = 0, kSlotMeanRn = 1, kSlotMeanX = 2;
int slotB(int l) { return 8 + 4 * l; }
int slotX(int l, int buf) { return 8 + 4 * l + 1 + buf; }
// Address of ctl->slot[i] computed without touching device memory.
double* slotPtr(SolveCtl* ctl, int i) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(ctl) + offsetof(SolveCtl, slot)) + i;
}
}  // namespace

PoissonDev::PoissonDev(nbl_ctx* ctx, const HostGrid& g, int kind, double tol, int maxIters,
                       const nbl_solver_opts& opts, int problems)
    : ctx_(ctx), kind_(kind), tol_(tol), maxIt_(maxIters), opts_(opts) {
  if (problems < 0) throw std::invalid_argument("a batch holds at least one problem");
  batched_ = problems > 0;
  nb_ = std::max(problems, 1);
  if (batched_) {
    if (kind != 0) throw std::invalid_argument("batched solves: mg-pcg only");
    if (opts.loops) throw std::invalid_argument("batched solves: the fused kernels only");
    if (opts.vcycle == 2) throw std::invalid_argument("batched solves: the one-kernel-per-sweep V-cycle only");
    if (opts.precision == NBL_PREC_MIXED_BF16) throw std::invalid_argument("batched solves: FP64 or FP32 levels");
    if (opts.launch_order == 2) throw std::invalid_argument("batched solves: plain order or bands of rows");
  }
  if (kind != 0 && kind != 1) throw std::invalid_argument("poisson kind must be 0 (mg-pcg) or 1 (cg-jacobi)");
  if (maxIters < 0) throw std::invalid_argument("max iterations must be >= 0");
  if (opts.pre_smooth < 1 || opts.post_smooth < 1)
    throw std::invalid_argument("pre/post smoothing sweeps must be >= 1");
  if (opts.precision < NBL_PREC_FP64 || opts.precision > NBL_PREC_MIXED_BF16)
    throw std::invalid_argument("unknown precision");
  singular_ = !hasOutlet(g);
  mixed_ = kind == 0 && opts.precision != NBL_PREC_FP64;
  bf16_ = kind == 0 && opts.precision == NBL_PREC_MIXED_BF16;
  if (opts.loops != 0 && opts.loops != 1) throw std::invalid_argument("loops must be 0 or 1");
  if (opts.loops && (kind != 0 || mixed_)) throw std::invalid_argument("loops: FP64 mg-pcg only");
  loops_ = opts.loops == 1;
  hl_ = buildLevels(g, kind == 0);
  // the null-space projections are no-ops without fluid cells
  project_ = singular_ && hl_[0].nf > 0;
  const int nl = int(hl_.size());
  const size_t ts = mixed_ ? sizeof(float) : sizeof(double);
  int stride = 1;
  for (int l = 0; l < nl; ++l) {
    const HostLevel& H = hl_[l];
    // a batch's problems share the grid: one copy of the codes per problem, so
    // that every cell kernel indexes codes and vectors alike
    const size_t cb = H.code.size() * sizeof(uint16_t);
    codes_.emplace_back(ctx, cb * size_t(nb_));
    for (int q = 0; q < nb_; ++q)
      NBL_HIP(hipMemcpy(codes_.back().as<char>() + size_t(q) * cb, H.code.data(), cb, hipMemcpyHostToDevice));
    MgLevel m{};
    m.nx = H.g.nx;
    m.ny = H.g.ny;
    m.nz = H.g.nz;
    m.n = (long long)H.g.cells();
    m.h2[0] = 1.0 / (H.g.dx * H.g.dx);
    m.h2[1] = 1.0 / (H.g.dy * H.g.dy);
    m.h2[2] = 1.0 / (H.g.dz * H.g.dz);
    m.code = codes_.back().as<uint16_t>();
    m.fx = H.f[0];
    m.fy = H.f[1];
    m.fz = H.f[2];
    m.nf = H.nf;
    // cells per thread along x (vector loads of each group).  Automatic: 2
    // where a row splits into at least 16 pairs; 4 on FP32 levels of 4M cells
    // and more (measured on 256^3: FP32 sweep 42 us at 4 vs 58 us at 2, FP64
    // sweep 99 us at 2 vs 102 us at 4; on 192x64x64, 2 is best for both)
    const int want = opts.cells_per_thread;
    int autoVec = (m.nx % 2 == 0 && m.nx >= 32) ? 2 : 1;
    if (mixed_ && m.nx % 4 == 0 && m.nx >= 64 && m.n >= (1LL << 22)) autoVec = 4;
    m.vec = want == 0 ? autoVec : (m.nx % want == 0 ? want : 1);
    m.slab = opts.launch_order == 0 ? slabFor(m.nx, m.ny, m.nz) : opts.launch_order - 1;
    if (batched_ && m.slab == 1) m.slab = 0;  // a batch's problems follow each other along the launch's z
    m.nb = nb_;
    dl_.push_back(m);
    // partials: per problem (the finalizes' counts); the buffer holds the batch's
    const int lb = levelGeo(m).blocks;
    blocks_.push_back(lb / nb_);
    stride = std::max(stride, lb);
    tblocks_.push_back(batched_ ? 0 : tileBlocks(m));
    stride = std::max(stride, tblocks_.back());
    const int rb = l > 0 ? restrictGeo(dl_[l - 1], m).blocks : 0;
    rblocks_.push_back(rb / nb_);
    stride = std::max(stride, rb);
    const size_t n = H.g.cells() * size_t(nb_);
    lb_.emplace_back(ctx, l > 0 ? n * ts : 0);
    lx0_.emplace_back(ctx, kind == 0 ? n * ts : 0);
    lx1_.emplace_back(ctx, kind == 0 ? n * ts : 0);
  }
  if (size_t(hl_[0].g.cells()) * size_t(nb_) >= (size_t(1) << 31))
    throw std::invalid_argument("grid too large: cell indices of the solve (all problems) must fit in 32 bits");
  if (kind == 0) coarseScr_ = DevBuf(ctx, 4 * hl_.back().g.cells() * ts);
  if (opts.coarse != NBL_COARSE_DIRECT && opts.coarse != NBL_COARSE_CG)
    throw std::invalid_argument("unknown coarse solve kind");
  if (opts.cells_per_thread != 0 && opts.cells_per_thread != 1 && opts.cells_per_thread != 2 &&
      opts.cells_per_thread != 4)
    throw std::invalid_argument("cells_per_thread must be 0, 1, 2 or 4");
  if (opts.launch_order < 0 || opts.launch_order > 3) throw std::invalid_argument("launch_order must be 0, 1, 2 or 3");
  if (opts.iteration_graph < 0 || opts.iteration_graph > 2)
    throw std::invalid_argument("iteration_graph must be 0, 1 or 2");
  const size_t n0 = hl_[0].g.cells();
  stride = std::max(stride, loopBlocks((long long)n0));
  if (loops_)
    for (int l = 0; l < nl; ++l) {
      ldiag_.emplace_back(ctx, hl_[l].g.cells() * 8);
      launch_diag(dl_[l], ldiag_.back().as<double>(), nullptr);
    }
  const size_t nv = n0 * size_t(nb_);  // the FP64 flexible-CG vectors of every problem
  if (mixed_) b0f_ = DevBuf(ctx, nv * sizeof(float));
  x_ = DevBuf(ctx, nv * 8);
  rhs_ = DevBuf(ctx, nv * 8);
  r_ = DevBuf(ctx, nv * 8);
  rn_ = DevBuf(ctx, nv * 8);
  z_ = DevBuf(ctx, nv * 8);
  p_ = DevBuf(ctx, nv * 8);
  p2_ = DevBuf(ctx, nv * 8);
  Ap_ = DevBuf(ctx, nv * 8);
  ctl_ = DevBuf(ctx, sizeof(SolveCtl) * size_t(nb_));
  redStride_ = stride;
  red_ = DevBuf(ctx, sizeof(double) * (kMaxRedVals + kSlotRows) * size_t(stride));
  NBL_HIP(hipMemset(ctl_.as<void>(), 0, sizeof(SolveCtl) * size_t(nb_)));
  const size_t nh = std::max<size_t>(2, size_t(nb_));
  NBL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hctl_), sizeof(SolveCtl) * nh, hipHostMallocDefault));
  std::memset(hctl_, 0, sizeof(SolveCtl) * nh);
  NBL_HIP(hipEventCreate(&e0_));
  NBL_HIP(hipEventCreate(&e1_));
  predicted_ = std::max(1, opts.check_every);
  if (kind == 0 && nl > 1 && opts.coarse == NBL_COARSE_DIRECT && !loops_) direct_ = buildDirectCoarse();
  if (batched_ && nl < 2) throw std::invalid_argument("batched solves: grids of at least two multigrid levels");
  if (batched_ && !direct_)
    throw std::invalid_argument("batched solves: the coarsest level must be solved directly (coarse DIRECT, "
                                "at most 4096 cells, an invertible operator)");
  if (opts.vcycle < 0 || opts.vcycle > 2) throw std::invalid_argument("vcycle must be 0, 1 or 2");
  // the two-kernels-per-level cycle: V(2,2), a level between the finest and
  // the directly solved coarsest one, and few enough level-0 workgroups
  // (kTiledFoldMax) that every reduction folds inline into its consumer.
  // Automatic: only where the BF16 smoother scratch asks for it.  Since the
  // per-sweep cycle dropped the singular systems' per-sweep projections it is
  // the faster one everywhere measured (periodic boxes, FP64 ms per step:
  // 64^3 3.42 vs 3.53, 100^3 9.12 vs 10.62, 128^3 9.86 vs 13.03); on the
  // 192x64x64 channel the tiled level-0 kernels, which recompute their halos,
  // take 33 + 24 us against 29 + 15 us for the per-sweep kernels they replace.
  const bool wantTiled = opts.vcycle == 2 || (opts.vcycle == 0 && bf16_);
  tiled_ = kind == 0 && !loops_ && !batched_ && wantTiled && nl >= 3 && opts.pre_smooth == 2 && opts.post_smooth == 2 &&
           direct_ && blocks_[0] <= kTiledFoldMax && tblocks_[0] <= kTiledFoldMax;
  for (int l = 0; tiled_ && l + 1 < nl; ++l) tiled_ = tileFits(dl_[l], dl_[l + 1]);
  if (tiled_ && project_) rnRaw_ = DevBuf(ctx, n0 * 8);
  // BF16 smoother scratch lives in the two-kernels-per-level cycle: the finest
  // level's pre-smoothed iterate passes from its down to its up kernel as BF16
  if (bf16_ && !tiled_)
    throw std::invalid_argument(
        "precision NBL_PREC_MIXED_BF16 needs the two-kernels-per-level V-cycle (vcycle 0 or 2, pre = post = 2, "
        "at least three levels, a direct coarsest solve, a finest level of at most 4096 tiles)");
  if (bf16_) xaB_ = DevBuf(ctx, n0 * sizeof(__hip_bfloat16));
  NBL_HIP(hipDeviceSynchronize());
}

// Dense inverse of the coarsest level's operator (multigrid.cpp:178-183 runs
// CgPoisson to 1e-6 there; an exact solve gives the same flexible-CG iteration
// counts and agrees with it far inside the outer tolerance).  Rows follow
// applyRow: fluid rows carry the 7-point couplings decoded from the level's
// face codes (periodic wraps included), solid rows are identity.  A singular
// system gets alpha*11^T on the fluid block and the inverse is projected
// afterwards.  Returns false (CG is used instead) when a pivot vanishes, e.g.
// for a fluid region the outlet does not reach.
bool PoissonDev::buildDirectCoarse() {
  const HostLevel& H = hl_.back();
  const long long nn = (long long)H.g.cells();
  if (nn > kDirectCoarseMax) return false;
  const int n = int(nn);
  const int dims[3] = {H.g.nx, H.g.ny, H.g.nz};
  const double h2[3] = {1.0 / (H.g.dx * H.g.dx), 1.0 / (H.g.dy * H.g.dy), 1.0 / (H.g.dz * H.g.dz)};
  const size_t w = size_t(2) * n;
  std::vector<double> G(size_t(n) * w, 0.0);
  double dmax = 0.0, dfl = 0.0;
  for (int c = 0; c < n; ++c) {
    double* row = G.data() + size_t(c) * w;
    row[n + c] = 1.0;
    const uint16_t code = H.code[size_t(c)];
    if (!(code & kFluidBit)) {
      row[c] = 1.0;
      continue;
    }
    const int q0[3] = {c % dims[0], (c / dims[0]) % dims[1], c / (dims[0] * dims[1])};
    double d = 0.0;
    for (int dir = 0; dir < 6; ++dir) {
      const int kd = kindOf(code, dir), a = dir >> 1;
      if (kd == 2) d += 2.0 * h2[a];
      if (kd != 1) continue;
      d += h2[a];
      int q[3] = {q0[0], q0[1], q0[2]};
      q[a] += (dir & 1) ? 1 : -1;
      q[a] = (q[a] + dims[a]) % dims[a];
      row[q[0] + dims[0] * (q[1] + dims[1] * q[2])] -= h2[a];
    }
    row[c] += d;
    dmax = std::max(dmax, d);
    dfl += d;
  }
  if (project_) {
    const double alpha = dfl / double(H.nf) / double(H.nf);
    for (int i = 0; i < n; ++i) {
      if (!(H.code[size_t(i)] & kFluidBit)) continue;
      double* row = G.data() + size_t(i) * w;
      for (int j = 0; j < n; ++j)
        if (H.code[size_t(j)] & kFluidBit) row[j] += alpha;
    }
  }
  hipStream_t s = nullptr;
  DevBuf g(ctx_, G.size() * sizeof(double)), col(ctx_, size_t(n) * 8), rowk(ctx_, w * 8), fail(ctx_, 256);
  NBL_HIP(hipMemcpy(g.as<void>(), G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
  NBL_HIP(hipMemset(fail.as<void>(), 0, 256));
  const double tiny = 1e-9 * std::max(dmax, 1.0);
  for (int k = 0; k < n; ++k) launch_gj_step(g.as<double>(), n, k, col.as<double>(), rowk.as<double>(), fail.as<int>(), tiny, s);
  int failed = 0;
  NBL_HIP(hipMemcpy(&failed, fail.as<void>(), sizeof(int), hipMemcpyDeviceToHost));
  if (failed) return false;
  coarseM_ = DevBuf(ctx_, size_t(n) * n * sizeof(double));
  coarseMf_ = DevBuf(ctx_, size_t(n) * n * sizeof(float));
  DevBuf scratch(ctx_, size_t(n + 1) * sizeof(double));
  launch_inv_store(g.as<double>(), n, dl_.back().code, H.nf, project_, scratch.as<double>(), coarseM_.as<double>(),
                   coarseMf_.as<float>(), s);
  NBL_HIP(hipDeviceSynchronize());
  return true;
}

PoissonDev::~PoissonDev() {
  if (iterGraph_) (void)hipGraphExecDestroy(iterGraph_);
  if (hctl_) (void)hipHostFree(hctl_);
  if (e0_) (void)hipEventDestroy(e0_);
  if (e1_) (void)hipEventDestroy(e1_);
}

MeanRef PoissonDev::meanRef(int slot, bool inl, int blocks, long long nf) const {
  if (!inl) return slotMean(slotPtr(ctl_.as<SolveCtl>(), slot));
  MeanRef m;
  m.parts = red_.as<double>() + size_t(kMaxRedVals + slot) * size_t(redStride_);
  m.n = blocks;
  m.nf = double(nf);
  return m;
}

MeanRef PoissonDev::zMeanTiled() const {
  MeanRef m;
  m.parts = slotRow(kSlotMeanZ);
  m.n = tblocks_[0];
  m.nf = double(dl_[0].nf);
  return m;
}

MeanRef PoissonDev::zMean() const {
  if (!project_) return MeanRef{};
  if (tiled_) return zMeanTiled();
  return meanRef(kSlotMeanZ, inlineMean(blocks_[0]), blocks_[0], dl_[0].nf);
}

RedBuf PoissonDev::meanRed(int slot, bool inl) const {
  if (!inl) return red();
  return RedBuf{red_.as<double>() + size_t(kMaxRedVals + slot) * size_t(redStride_), redStride_};
}

void PoissonDev::meanFin(int slot, bool inl, int blocks, long long nf, bool predicated, hipStream_t s) {
  if (!inl) fin(FIN_MEAN, 1, blocks, slot, nf, predicated, s);
}

void PoissonDev::fin(int op, int K, int nparts, int slot, long long nf, bool predicated, hipStream_t s,
                     int earlyExit, double convTol, int row0) {
  FinArgs a{};
  a.op = op;
  a.K = K;
  a.nparts = nparts;
  a.slot = slot;
  a.nf = nf;
  a.convTol = convTol;
  a.earlyExit = earlyExit;
  a.predicated = predicated ? 1 : 0;
  const RedBuf rb = red();
  launch_fin(RedBuf{rb.partials + size_t(row0) * size_t(rb.stride), rb.stride}, a, ctl_.as<SolveCtl>(), s, nb_);
}

// In-situ timing of the finest level's plain sweeps (the roofline kernel):
// each timed launch gets a slice of the stamp buffer in which every block
// records the device wall clock when it starts and when all its waves are
// done (no events or extra packets in the stream, and no fence: the solve
// runs as untimed).  A launch lasts from its first block start to its last block end;
// launches a converged solve skips on the device leave their slice at zero.
long long* PoissonDev::timedSweep(int l, bool plain) {
  if (batched_ || !timeSweeps_ || l != 0 || !plain || stampUsed_ >= stampCap_) return nullptr;
  return stamps_.as<long long>() + size_t(stampUsed_++) * 2 * size_t(blocks_[0]);
}

void PoissonDev::setSweepTiming(bool on) {
  timeSweeps_ = on;
  if (!on) return;
  sweepMs_ = 0.0;
  sweepN_ = 0;
  stampUsed_ = 0;
  if (!stamps_.bytes()) {
    stampCap_ = kStampLaunches;
    stamps_ = DevBuf(ctx_, size_t(stampCap_) * 2 * size_t(blocks_[0]) * sizeof(long long));
    int dev = 0, khz = 0;
    NBL_HIP(hipGetDevice(&dev));
    NBL_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    wallKHz_ = khz > 0 ? khz : 100000;
  }
  NBL_HIP(hipMemset(stamps_.as<void>(), 0, stamps_.bytes()));
}

void PoissonDev::sweepTiming(double* ms, long long* n, double* bytes) {
  if (stampUsed_ > 0) {
    NBL_HIP(hipDeviceSynchronize());
    const size_t nb = size_t(blocks_[0]);
    std::vector<long long> h(size_t(stampUsed_) * 2 * nb);
    NBL_HIP(hipMemcpy(h.data(), stamps_.as<void>(), h.size() * sizeof(long long), hipMemcpyDeviceToHost));
    for (int q = 0; q < stampUsed_; ++q) {
      const long long* t = h.data() + size_t(q) * 2 * nb;
      long long t0 = t[0], t1 = t[1];
      bool ran = true;
      for (size_t b = 0; b < nb && ran; ++b) {
        if (t[2 * b] == 0 || t[2 * b + 1] == 0) ran = false;
        t0 = std::min(t0, t[2 * b]);
        t1 = std::max(t1, t[2 * b + 1]);
      }
      if (!ran) continue;
      sweepMs_ += double(t1 - t0) / double(wallKHz_);
      ++sweepN_;
    }
    NBL_HIP(hipMemset(stamps_.as<void>(), 0, size_t(stampUsed_) * 2 * nb * sizeof(long long)));
    stampUsed_ = 0;
  }
  if (ms) *ms = sweepN_ ? sweepMs_ / double(sweepN_) : 0.0;
  if (n) *n = sweepN_;
  if (bytes) *bytes = double(dl_[0].n) * (3.0 * (mixed_ ? 4.0 : 8.0) + 2.0);
}

// One V(pre, post) cycle from level l (multigrid.cpp:185-198).  Each sweep and
// each restriction leaves its null-space mean pending in a slot; consumers
// subtract it on read.
template <class T>
void PoissonDev::vcycle(int l, const T* b, MeanRef mb, const T** xout, MeanRef* mxout, hipStream_t s) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[l];
  const RedBuf rb = red();
  if (l + 1 == nlevels()) {
    T* x = lx0_[l].as<T>();
    if (direct_)
      launch_mg_coarse_direct<T>(L, coarseM_.as<double>(), coarseMf_.as<float>(), b, mb, x, ctl, s);
    else  // the one-workgroup CG reads its right-hand side mean from SolveCtl
      launch_mg_coarse<T>(L, project_, b, mb.slot, x, coarseScr_.as<T>(), ctl, s);
    *xout = x;
    *mxout = MeanRef{};
    return;
  }
  T* bufs[2] = {lx0_[l].as<T>(), lx1_[l].as<T>()};
  const int nbl = blocks_[l];
  const bool inlX = inlineMean(nbl);
  // Singular systems: the reference removes the fluid mean after every sweep
  // and of every restricted residual (multigrid.cpp:110-112, 150-152).  The
  // operator annihilates constants on fluid cells and is symmetric, so a
  // residual's fluid sum vanishes and those projections change neither a
  // residual nor, after the final projection of z (launch_mg_finish), the
  // cycle's output in exact arithmetic.  The cycle keeps only the projection
  // of z and, where the coarsest level is solved by CG, of its right-hand side
  // (the projected inverse of the direct solve ignores a mean); results agree
  // with the reference's to rounding, and the fusions of the non-singular
  // cycle (first sweeps from zero written by the restrictions) apply.
  const bool sp = false;
  const bool rp = project_ && l + 2 == nlevels() && !direct_;
  const MeanRef ms[2] = {sp ? meanRef(slotX(l, 0), inlX, nbl, L.nf) : MeanRef{},
                         sp ? meanRef(slotX(l, 1), inlX, nbl, L.nf) : MeanRef{}};
  auto xRed = [&](int o) { return sp ? meanRed(slotX(l, o), inlX) : rb; };
  int cur = -1;
  int q0 = 0;
  const bool fusedL0 = l == 0 && l0Fused_;
  // Without a pending mean on b the first sweep from zero is pointwise: the
  // level-0 update (or r_new's projection) and each restriction write it with
  // the right-hand side.
  const bool firstDone = opts_.pre_smooth >= 1 && ((l == 0 && l0First_) || (l > 0 && !mb.on()));
  if (firstDone) {
    cur = 0;
    q0 = 1;
  }
  for (int q = q0; q < opts_.pre_smooth; ++q) {
    const int o = cur == 0 ? 1 : 0;
    InFold conv{};
    if (l == 0 && conv_.n && cur < 0) {
      // the flexible CG's convergence test rides on the first sweep from zero
      // (it reads no pending mean; the kernel issues its loads before the test)
      if (mb.on()) throw std::logic_error("convergence sweep with a pending mean");
      conv = conv_;
      conv_ = InFold{};
    }
    launch_mg_sweep<T>(L, b, mb, cur < 0 ? nullptr : bufs[cur], cur < 0 ? MeanRef{} : ms[cur], cur < 0, nullptr,
                       MeanRef{}, nullptr, bufs[o], sp, xRed(o), ctl, s, timedSweep(l, cur >= 0), conv);
    if (sp) meanFin(slotX(l, o), inlX, nbl, L.nf, true, s);
    cur = o;
  }
  const MgLevel& Lc = dl_[l + 1];
  T* bc = lb_[l + 1].as<T>();
  const int rbl = rblocks_[l + 1];
  const bool inlB = inlineMean(rbl) && !(l + 2 == nlevels() && !direct_);
  const MeanRef mbc = rp ? meanRef(slotB(l + 1), inlB, rbl, Lc.nf) : MeanRef{};
  // the coarse level's first sweep, unless the coarse level is the coarsest
  T* x1c = (!rp && opts_.pre_smooth >= 1 && l + 2 < nlevels()) ? lx0_[l + 1].as<T>() : nullptr;
  // Otherwise the test rides on the level-0 restriction (the first level-0
  // sweep of the fused path is then the plain smoother sweep, kept free of it).
  InFold conv{};
  if (l == 0 && conv_.n) {
    if (mb.on() || ms[cur].on()) throw std::logic_error("convergence restriction with a pending mean");
    conv = conv_;
    conv_ = InFold{};
  }
  launch_mg_resid_restrict<T>(L, Lc, b, mb, bufs[cur], ms[cur], bc, rp, rp ? meanRed(slotB(l + 1), inlB) : rb,
                              ctl, s, x1c, conv);
  if (rp) meanFin(slotB(l + 1), inlB, rbl, Lc.nf, true, s);
  const T* ec = nullptr;
  MeanRef mec{};
  vcycle<T>(l + 1, bc, mbc, &ec, &mec, s);
  for (int q = 0; q < opts_.post_smooth; ++q) {
    const int o = 1 - cur;
    if (fusedL0 && q == opts_.post_smooth - 1 && q > 0) {
      // last sweep: writes z and the flexible-CG dots instead of x
      const bool zs = project_;  // (only with zDotsMean(): minv fuses singular dots then only)
      launch_mg_sweep_dots<T>(L, b, bufs[cur], z_.as<double>(), l0Rn_, l0R_,
                              zs ? RedBuf{rb.partials + kZDotsRow * size_t(rb.stride), rb.stride} : rb, ctl, s,
                              inline_, zs);
      *xout = nullptr;
      *mxout = MeanRef{};
      return;
    }
    // the last level-0 sweep folds its fluid sum where the dots kernel removes
    // z's mean (l0FoldLast_)
    const bool fl = sp || (l == 0 && l0FoldLast_ && q == opts_.post_smooth - 1);
    launch_mg_sweep<T>(L, b, mb, bufs[cur], ms[cur], false, q == 0 ? ec : nullptr, q == 0 ? mec : MeanRef{}, &Lc,
                       bufs[o], fl, fl ? meanRed(slotX(l, o), inlX) : rb, ctl, s, timedSweep(l, q > 0));
    if (fl) meanFin(slotX(l, o), inlX, nbl, L.nf, true, s);
    cur = o;
    if (fl && !sp) {
      *xout = bufs[cur];
      *mxout = meanRef(slotX(l, cur), inlX, nbl, L.nf);
      return;
    }
  }
  *xout = bufs[cur];
  *mxout = ms[cur];
}

// z = Minv r: one V-cycle from a zero guess, then the level-0 projection
// (multigrid.cpp:283-289).  z's mean is left pending in kSlotMeanZ (zMean()).  With
// fuseDots (systems without projections) the V-cycle output is left for
// k_fcg_dots_x, which forms z itself.
void PoissonDev::minv(const double* r, hipStream_t s, const double* rOld, bool fused, bool b0Ready, bool x1Ready) {
  if (tiled_) {  // the first cycle of a solve, or precondition: z = Minv r, z's mean left pending
    if (rOld || fused) throw std::logic_error("minv: the tiled cycle forms iterations itself");
    if (mixed_) {
      launch_to_level0<float>(dl_[0], r, b0f_.as<float>(), ctl_.as<SolveCtl>(), s);
      tiledCycle<float>(s, 0, true, b0f_.as<float>(), nullptr, nullptr, nullptr, InFold{});
    } else {
      tiledCycle<double>(s, 0, true, r, nullptr, nullptr, nullptr, InFold{});
    }
    ++vcount_;
    return;
  }
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[0];
  const RedBuf rb = red();
  double* z = z_.as<double>();
  const bool dots = rOld != nullptr;
  if (fused && project_ && !zDotsMean()) throw std::logic_error("minv: singular fused dots need inline folds");
  l0Fused_ = fused;
  l0First_ = fused || x1Ready;
  l0Rn_ = r;
  l0R_ = rOld;
  const bool inlZ = inlineMean(blocks_[0]);
  const RedBuf zRed = project_ ? meanRed(kSlotMeanZ, inlZ) : rb;
  // singular systems with the flexible-CG dots: the last level-0 sweep folds
  // its fluid sum and the dots kernel forms z = x - mean (no k_finish)
  l0FoldLast_ = project_ && dots && !fused;
  // z from the cycle's output x: with the dots in one pass where x carries
  // its pending mean (or needs none); otherwise a pass writing z (and, on
  // singular systems, its fluid sum, a single-level cycle's output having
  // none folded) before the dots
  auto finish = [&](const auto* x, MeanRef mx) {
    using T = std::remove_cv_t<std::remove_pointer_t<decltype(x)>>;
    if (!x) return;
    if (dots && (!project_ || mx.on())) {
      launch_fcg_dots_x<T>(L, x, mx, z, r, rOld, ctl, rb, s, project_ && inline_);
      return;
    }
    launch_mg_finish<T>(L, x, mx, z, project_, zRed, ctl, s);
    if (project_) meanFin(kSlotMeanZ, inlZ, blocks_[0], L.nf, true, s);
    if (dots) launch_fcg_dots(L, z, zMean(), r, rOld, ctl, rb, s, inline_);
  };
  if (mixed_) {
    if (!fused && !b0Ready) launch_to_level0<float>(L, r, b0f_.as<float>(), ctl, s);
    const float* x = nullptr;
    MeanRef mx{};
    vcycle<float>(0, b0f_.as<float>(), MeanRef{}, &x, &mx, s);
    finish(x, mx);
  } else {
    const double* x = nullptr;
    MeanRef mx{};
    vcycle<double>(0, r, MeanRef{}, &x, &mx, s);
    finish(x, mx);
  }
  l0Fused_ = false;
  l0First_ = false;
  l0FoldLast_ = false;
  ++vcount_;
}

// One iteration of the flexible CG loop (multigrid.cpp:309-340) or of the
// Jacobi CG loop (poisson.cpp:197-225); every launch is a no-op once the
// device has set ctl->done.  Iteration k >= 1 forms p = z + beta p of the
// previous iteration inside its A p, so p alternates between two buffers,
// and the flexible CG's r / r_new alternate instead of being copied.
// With inline_ the three finalizes of an iteration run inside their
// consumers: beta in the next A p, alpha in the update, the convergence test
// in the first level-0 sweep; the partial rows then alternate so that no
// kernel overwrites the rows it folds (z.r_new, z.r: rows 0-1; p.Ap: row 2;
// |r_new|^2: row 3).
void PoissonDev::enqueueIteration(hipStream_t s, int k) {
  if (loops_) {
    enqueueIterationLoops(s);
    return;
  }
  if (tiled_) {
    enqueueIterationTiled(s, k);
    return;
  }
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[0];
  const RedBuf rb = red();
  const int nb = blocks_[0];
  const RedBuf rbP = inline_ ? RedBuf{rb.partials + 2 * size_t(rb.stride), rb.stride} : rb;
  const RedBuf rbR = inline_ ? RedBuf{rb.partials + 3 * size_t(rb.stride), rb.stride} : rb;
  auto fold = [&](const RedBuf& r) { return inline_ ? InFold{r.partials, r.stride, nb, k} : InFold{}; };
  double* x = x_.as<double>();
  double* z = z_.as<double>();
  double* Ap = Ap_.as<double>();
  double* p = k == 0 ? p_.as<double>() : ((k & 1) ? p2_.as<double>() : p_.as<double>());
  if (k == 0) {
    launch_fcg_apply(L, p, Ap, ctl, rbP, s);
  } else {
    const double* pOld = (k & 1) ? p_.as<double>() : p2_.as<double>();
    if (zDotsMean() && inline_)
      launch_fcg_apply_pu(L, z, pOld, p, Ap, ctl, rbP, s,
                          InFold{rb.partials + kZDotsRow * size_t(rb.stride), rb.stride, nb, k}, double(L.nf));
    else if (zDotsMean())  // beta and z's mean from FIN_BETA_FLEX_MEAN
      launch_fcg_apply_pu(L, z, pOld, p, Ap, ctl, rbP, s, InFold{}, double(L.nf), slotPtr(ctl, kSlotMeanZ));
    else
      launch_fcg_apply_pu(L, z, pOld, p, Ap, ctl, rbP, s, fold(rb));
  }
  if (!inline_) fin(FIN_ALPHA, 1, nb, 0, 0, true, s);
  if (kind_ == 0) {
    double* r = (k & 1) ? rn_.as<double>() : r_.as<double>();
    double* rn = (k & 1) ? r_.as<double>() : rn_.as<double>();
    if (project_) {
      const bool first = nlevels() > 1 && opts_.pre_smooth >= 1;
      if (first) {
        // The right-hand side is projected once at the start and A annihilates
        // constants on fluid cells and is symmetric, so r - alpha Ap stays
        // mean-free up to rounding: r_new needs no projection of its own, and
        // the update writes |r_new|^2, the mixed form's FP32 copy and the
        // level-0 first sweep from zero as on non-singular systems.
        if (mixed_)
          launch_fcg_update_first<float>(L, x, p, r, Ap, rn, lx0_[0].as<float>(), b0f_.as<float>(), ctl, rbR, s,
                                         fold(rbP));
        else
          launch_fcg_update_first<double>(L, x, p, r, Ap, rn, lx0_[0].as<double>(), nullptr, ctl, rbR, s,
                                          fold(rbP));
      } else {
        // a single-level cycle: r_new's projection rides on the norm kernel
        const bool inlM = inlineMean(nb);
        launch_fcg_update(L, true, x, p, r, Ap, rn, ctl, meanRed(kSlotMeanRn, inlM), s, fold(rbP));
        meanFin(kSlotMeanRn, inlM, nb, L.nf, true, s);
        launch_fcg_norm(L, rn, meanRef(kSlotMeanRn, inlM, nb, L.nf), ctl, rbR, s,
                        mixed_ ? b0f_.as<float>() : nullptr, nullptr, nullptr);
      }
      if (inline_) conv_ = fold(rbR);
      else fin(FIN_CONV, 1, nb, 0, 0, true, s);
      // the last level-0 sweep writes z and the dots, z's mean left for the
      // next A p (zDotsMean); otherwise k_fcg_dots_x
      minv(rn, s, r, zDotsMean(), mixed_, first);
      if (conv_.n) throw std::logic_error("flexible CG: convergence test not placed");
    } else {
      // fusions need a first pre-sweep and a coarser level to smooth above
      const bool fused = nlevels() > 1 && opts_.pre_smooth >= 1;
      if (fused && mixed_)
        launch_fcg_update_first<float>(L, x, p, r, Ap, rn, lx0_[0].as<float>(), b0f_.as<float>(), ctl, rbR, s,
                                       fold(rbP));
      else if (fused)
        launch_fcg_update_first<double>(L, x, p, r, Ap, rn, lx0_[0].as<double>(), nullptr, ctl, rbR, s,
                                        fold(rbP));
      else
        launch_fcg_update(L, false, x, p, r, Ap, rn, ctl, rb, s);
      if (inline_) conv_ = fold(rbR);
      else fin(FIN_CONV, 1, nb, 0, 0, true, s);
      minv(rn, s, r, fused);
      if (conv_.n) throw std::logic_error("flexible CG: convergence test not placed");
    }
    if (!inline_ && zDotsMean())
      fin(FIN_BETA_FLEX_MEAN, 3, nb, kSlotMeanZ, L.nf, true, s, 0, 0.0, kZDotsRow);
    else if (!inline_)
      fin(FIN_BETA_FLEX, 2, nb, 0, 0, true, s);
  } else {
    double* r = r_.as<double>();
    launch_cg_update(L, x, p, r, Ap, ctl, rb, s);
    fin(FIN_CONV, 1, nb, 0, 0, true, s);
    launch_cg_z(L, r, z, false, nullptr, ctl, rb, s);
    fin(FIN_BETA, 1, nb, 0, 0, true, s);
  }
}

// One V(2,2) cycle in two kernels per level (vcycle.hip) and the coarsest
// solve between them.  Partial rows: |r_new|^2 in row 3 (the convergence
// test, folded by level 1's down kernel), the flexible-CG dots in rows 0-1,
// each level's restricted right-hand side's fluid sum in its slotB row, z's
// in kSlotMeanZ's.  Launches: (levels - 1) down, 1 coarsest, (levels - 1) up.
template <class T>
void PoissonDev::tiledCycle(hipStream_t s, int k, bool first, const T* b0, const double* r, double* rn,
                            const double* p, InFold alpha) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const int nl = nlevels();
  const RedBuf rb = red();
  double* rowNorm = rb.partials + 3 * size_t(rb.stride);
  auto meanOf = [&](int slot, int producerBlocks, long long nf) {
    if (!project_) return MeanRef{};
    MeanRef m;
    m.parts = slotRow(slot);
    m.n = producerBlocks;
    m.nf = double(nf);
    return m;
  };
  // level 0 down
  DownArgs<T> d0;
  d0.L = dl_[0];
  d0.Lc = dl_[1];
  d0.ctl = ctl;
  d0.xa = lx0_[0].as<T>();
  if (bf16_) d0.xaB = xaB_.as<__hip_bfloat16>();
  d0.bc = lb_[1].as<T>();
  d0.rowMean = project_ ? slotRow(slotB(1)) : nullptr;
  int src = DS_B;
  if (first) {
    d0.b = b0;
  } else {
    d0.rnOut = rn;
    d0.rowNorm = rowNorm;
    if constexpr (std::is_same_v<T, float>) d0.b0Out = b0f_.as<float>();
    if (project_) {
      src = DS_RN;
      d0.rnRaw = rnRaw_.as<double>();
      d0.mrn = meanOf(kSlotMeanRn, blocks_[0], dl_[0].nf);
    } else {
      src = DS_UPD;
      d0.r = r;
      d0.Ap = Ap_.as<double>();
      d0.x = x_.as<double>();
      d0.p = p;
      d0.alpha = alpha;
    }
  }
  launch_down<T>(src, d0, s);
  // levels 1 .. nl-2 down; level 1 carries the convergence test
  for (int l = 1; l + 1 < nl; ++l) {
    DownArgs<T> d;
    d.L = dl_[l];
    d.Lc = dl_[l + 1];
    d.ctl = ctl;
    d.b = lb_[l].as<T>();
    d.mb = meanOf(slotB(l), tblocks_[l - 1], dl_[l].nf);
    d.xa = lx0_[l].as<T>();
    d.bc = lb_[l + 1].as<T>();
    d.rowMean = project_ ? slotRow(slotB(l + 1)) : nullptr;
    if (l == 1 && !first) d.conv = InFold{rowNorm, rb.stride, tblocks_[0], k};
    launch_down<T>(DS_B, d, s);
  }
  const int lc = nl - 1;
  launch_mg_coarse_direct<T>(dl_[lc], coarseM_.as<double>(), coarseMf_.as<float>(), lb_[lc].as<T>(),
                             meanOf(slotB(lc), tblocks_[lc - 1], dl_[lc].nf), lx0_[lc].as<T>(), ctl, s);
  const T* e = lx0_[lc].as<T>();
  for (int l = nl - 2; l >= 1; --l) {
    UpArgs<T> u;
    u.L = dl_[l];
    u.Lc = dl_[l + 1];
    u.b = lb_[l].as<T>();
    u.mb = meanOf(slotB(l), tblocks_[l - 1], dl_[l].nf);
    u.xa = lx0_[l].as<T>();
    u.e = e;
    u.xb = lx1_[l].as<T>();
    u.ctl = ctl;
    launch_up<T>(UD_X, u, s);
    e = lx1_[l].as<T>();
  }
  UpArgs<T> u0;
  u0.L = dl_[0];
  u0.Lc = dl_[1];
  u0.b = b0;
  u0.xa = lx0_[0].as<T>();
  if (bf16_) u0.xaB = xaB_.as<__hip_bfloat16>();
  u0.e = e;
  u0.z = z_.as<double>();
  u0.ctl = ctl;
  if (!first && !project_) {
    u0.rn = rn;
    u0.r = r;
    u0.rb = rb;
    u0.countIt = 1;
    launch_up<T>(UD_DOTS, u0, s);
  } else {
    u0.rb = project_ ? RedBuf{slotRow(kSlotMeanZ), rb.stride} : RedBuf{nullptr, 0};
    launch_up<T>(UD_Z, u0, s);
  }
}

// One flexible-CG iteration (multigrid.cpp:309-340) around the tiled cycle:
// A p (with p = z + beta p of the previous iteration and beta folded in), the
// cycle (whose level-0 down kernel performs the update with alpha folded in,
// and whose level-1 down kernel the convergence test), and on singular
// systems the update writing r_new's fluid sums first and the dots after.
void PoissonDev::enqueueIterationTiled(hipStream_t s, int k) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[0];
  const RedBuf rb = red();
  const int nb = blocks_[0];
  const RedBuf rbP{rb.partials + 2 * size_t(rb.stride), rb.stride};
  double* x = x_.as<double>();
  double* z = z_.as<double>();
  double* Ap = Ap_.as<double>();
  double* p = k == 0 ? p_.as<double>() : ((k & 1) ? p2_.as<double>() : p_.as<double>());
  if (k == 0) {
    launch_fcg_apply(L, p, Ap, ctl, rbP, s);
  } else {
    const double* pOld = (k & 1) ? p_.as<double>() : p2_.as<double>();
    // the dots come from the level-0 up kernel, or on singular systems from k_fcg_dots
    launch_fcg_apply_pu(L, z, pOld, p, Ap, ctl, rbP, s, InFold{rb.partials, rb.stride, project_ ? nb : tblocks_[0], k});
  }
  const InFold alpha{rbP.partials, rbP.stride, nb, k};
  double* r = (k & 1) ? rn_.as<double>() : r_.as<double>();
  double* rn = (k & 1) ? r_.as<double>() : rn_.as<double>();
  if (project_)
    launch_fcg_update(L, true, x, p, r, Ap, rnRaw_.as<double>(), ctl, meanRed(kSlotMeanRn, true), s, alpha);
  if (mixed_) tiledCycle<float>(s, k, false, b0f_.as<float>(), r, rn, p, alpha);
  else tiledCycle<double>(s, k, false, rn, r, rn, p, alpha);
  if (project_) launch_fcg_dots(L, z, zMeanTiled(), rn, r, ctl, rb, s, true);
  ++vcount_;
}

// ---------------------------------------------------------------------------
// The reference's loop structure, one kernel per loop (opts.loops): smooth,
// residual, restrictResidual, prolongAdd, vcycle with CgPoisson(1e-6, 500) on
// the coarsest level, removeMean, and the flexible-CG loop of
// MgPoisson::solve (multigrid.cpp:81-194, 283-340), each loop a launch and each
// dot product a reduction with its finalize.  Level buffers: x in lx0_, the
// right-hand side in lb_ (level 0: the caller's vector), A x / r in lx1_.
// ---------------------------------------------------------------------------
void PoissonDev::removeMeanLoops(int l, double* x, hipStream_t s) {
  if (!project_) return;
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[l];
  launch_loop_msum(L, x, red(), ctl, s);
  fin(FIN_MEAN, 1, loopBlocks(L.n), slotX(l, 0), L.nf, true, s);
  launch_loop_msub(L, x, slotPtr(ctl, slotX(l, 0)), ctl, s);
}

void PoissonDev::smoothLoops(int l, const double* b, double* x, int sweeps, hipStream_t s) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[l];
  double* Ax = lx1_[l].as<double>();
  for (int q = 0; q < sweeps; ++q) {
    launch_loop_apply(L, x, Ax, ctl, s);
    launch_loop_jacobi(L, b, Ax, ldiag_[l].as<double>(), x, ctl, s);
    removeMeanLoops(l, x, s);
  }
}

void PoissonDev::vcycleLoops(int l, const double* b, double* x, hipStream_t s) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[l];
  if (l + 1 == nlevels()) {
    launch_mg_coarse<double>(L, project_, b, nullptr, x, coarseScr_.as<double>(), ctl, s);
    return;
  }
  smoothLoops(l, b, x, opts_.pre_smooth, s);
  double* r = lx1_[l].as<double>();
  launch_loop_apply(L, x, r, ctl, s);
  launch_loop_resid(L, b, r, ctl, s);
  const MgLevel& Lc = dl_[l + 1];
  double* rc = lb_[l + 1].as<double>();
  launch_loop_restrict(L, Lc, r, rc, ctl, s);
  removeMeanLoops(l + 1, rc, s);
  double* ec = lx0_[l + 1].as<double>();
  launch_loop_copy(Lc.n, ec, nullptr, ctl, s);
  vcycleLoops(l + 1, rc, ec, s);
  launch_loop_prolong(L, Lc, ec, x, ctl, s);
  smoothLoops(l, b, x, opts_.post_smooth, s);
}

// Minv (multigrid.cpp:283-289): z = 0, one V-cycle, removeMean(z).
void PoissonDev::minvLoops(const double* r, hipStream_t s) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  double* z = z_.as<double>();
  launch_loop_copy(dl_[0].n, z, nullptr, ctl, s);
  vcycleLoops(0, r, z, s);
  removeMeanLoops(0, z, s);
  ++vcount_;
}

// One iteration of multigrid.cpp:309-340, one kernel per loop; r and r_new
// are copied like the reference's `r = rnew`.
void PoissonDev::enqueueIterationLoops(hipStream_t s) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[0];
  const RedBuf rb = red();
  const long long n = L.n;
  const int nb = loopBlocks(n);
  double *x = x_.as<double>(), *z = z_.as<double>(), *p = p_.as<double>(), *Ap = Ap_.as<double>();
  double *r = r_.as<double>(), *rn = rn_.as<double>();
  launch_loop_apply(L, p, Ap, ctl, s);
  launch_loop_dot(n, p, Ap, nullptr, rb, ctl, s);
  fin(FIN_ALPHA, 1, nb, 0, 0, true, s);
  launch_loop_update(n, x, p, r, Ap, rn, ctl, s);
  removeMeanLoops(0, rn, s);
  launch_loop_dot(n, rn, rn, nullptr, rb, ctl, s);
  fin(FIN_CONV, 1, nb, 0, 0, true, s);
  minvLoops(rn, s);
  launch_loop_dot(n, z, rn, r, rb, ctl, s);
  fin(FIN_BETA_FLEX, 2, nb, 0, 0, true, s);
  launch_loop_pupdate(n, p, z, ctl, s);
  launch_loop_copy(n, r, rn, ctl, s);
}

// z = Minv r alone (multigrid.cpp:283-289): one V-cycle from zero and the
// level-0 projection.  d_r and d_z are device vectors of the level-0 size.
void PoissonDev::precondition(const double* d_r, double* d_z, hipStream_t s) {
  if (kind_ != 0) throw std::invalid_argument("precondition: mg-pcg solver only");
  if (batched_) throw std::invalid_argument("precondition: not on a batch");
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[0];
  launch_ctl_init(ctl, tol_, s);
  if (loops_) minvLoops(d_r, s);
  else minv(d_r, s, nullptr, false);
  // z - pending mean (projection) into z, p; the partials it writes are unused
  launch_fcg_start(L, z_.as<double>(), loops_ ? MeanRef{} : zMean(), d_r, p_.as<double>(), ctl, red(), s);
  NBL_HIP(hipMemcpyAsync(d_z, z_.as<void>(), size_t(L.n) * 8, hipMemcpyDeviceToDevice, s));
}

bool PoissonDev::iterationPair(hipStream_t s) {
  if (opts_.iteration_graph != 2 || loops_ || timeSweeps_ || graphFailed_) return false;
  if (!iterGraph_) {
    const int v0 = vcount_;
    hipGraph_t g = nullptr;
    NBL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      enqueueIteration(s, 1);
      enqueueIteration(s, 2);
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      vcount_ = v0;
      graphFailed_ = true;
      throw;
    }
    NBL_HIP(hipStreamEndCapture(s, &g));
    graphV_ = vcount_ - v0;
    vcount_ = v0;
    const hipError_t e = hipGraphInstantiate(&iterGraph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      iterGraph_ = nullptr;
      graphFailed_ = true;
      return false;
    }
  }
  NBL_HIP(hipGraphLaunch(iterGraph_, s));
  vcount_ += graphV_;
  return true;
}

void PoissonDev::readCtl(hipStream_t s) {
  NBL_HIP(hipMemcpyAsync(hctl_, ctl_.as<void>(), sizeof(SolveCtl) * size_t(nb_), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
}

// MgPoisson::solve (multigrid.cpp:268-355) / CgPoisson::solve
// (poisson.cpp:136-252).  The host enqueues the iteration count the previous
// solve needed, then polls the device flag every check_every iterations.
nbl_poisson_report PoissonDev::solve(const double* d_b, double* d_x, hipStream_t s) {
  // the cell kernels read and write pairs of cells as 16-byte vectors
  if ((reinterpret_cast<uintptr_t>(d_b) | reinterpret_cast<uintptr_t>(d_x)) % 16)
    throw std::invalid_argument("poisson solve: device vectors must be 16-byte aligned");
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const MgLevel& L = dl_[0];
  const RedBuf rb = red();
  const int nb = blocks_[0];
  launch_ctl_init(ctl, tol_, s, nb_);
  NBL_HIP(hipEventRecord(e0_, s));
  vcount_ = 0;
  inline_ = kind_ == 0 && !loops_ && !batched_ && nlevels() > 1 && opts_.pre_smooth >= 2 &&
            opts_.post_smooth >= 2 && nb <= kInlineFoldMax;
  conv_ = InFold{};
  double* x = x_.as<double>();
  double* rhs = rhs_.as<double>();
  double* r = r_.as<double>();
  double* z = z_.as<double>();
  double* p = p_.as<double>();
  const bool mg = kind_ == 0;
  // MgPoisson zeroes the right-hand side on solid cells; CgPoisson does not.
  // The kernel that writes r also writes the first V-cycle's FP32 right-hand
  // side and level-0 first sweep from zero.
  const bool first = mg && !loops_ && !tiled_ && nlevels() > 1 && opts_.pre_smooth >= 1;
  float* b0 = first && mixed_ ? b0f_.as<float>() : nullptr;
  float* x1f = first && mixed_ ? lx0_[0].as<float>() : nullptr;
  double* x1d = first && !mixed_ ? lx0_[0].as<double>() : nullptr;
  if (project_) {
    launch_fcg_init(L, mg, d_b, rhs, x, rb, s);
    fin(FIN_MEAN, 1, nb, kSlotMeanRhs, L.nf, false, s);
    launch_fcg_init2(L, true, rhs, r, ctl, rb, s, b0, x1f, x1d);
  } else {
    launch_fcg_init(L, mg, d_b, rhs, x, rb, s, r, b0, x1f, x1d);
  }
  fin(FIN_BNORM, 1, nb, 0, 0, false, s, mg ? 1 : 0);
  if (mg) {
    if (loops_) minvLoops(r, s);
    else minv(r, s, nullptr, false, b0 != nullptr, first);
    launch_fcg_start(L, z, loops_ ? MeanRef{} : zMean(), r, p, ctl, rb, s);
  } else {
    launch_cg_z(L, r, z, true, p, ctl, rb, s);
  }
  fin(FIN_RZ, 1, nb, 0, 0, true, s);
  int enq = 0;
  int batch = std::clamp(predicted_, 1, std::max(1, maxIt_));
  const int step = std::max(1, opts_.check_every);
  while (enq < maxIt_) {
    for (int q = 0; q < batch && enq < maxIt_;) {
      if (enq % 2 == 1 && q + 2 <= batch && enq + 2 <= maxIt_ && iterationPair(s)) {
        q += 2;
        enq += 2;
      } else {
        enqueueIteration(s, enq);
        ++q;
        ++enq;
      }
    }
    readCtl(s);
    bool done = true;
    for (int q = 0; q < nb_; ++q) done = done && hctl_[q].done;
    if (done) break;
    batch = step;
  }
  if (project_) {
    launch_fluid_sum(L, x, rb, s);
    fin(FIN_MEAN, 1, nb, kSlotMeanX, L.nf, false, s);
  }
  launch_fcg_final(L, project_, x, rhs, d_x, ctl, rb, s);
  fin(FIN_RESID, 1, nb, 0, 0, false, s, 0, mg ? tol_ : 1e-6);
  NBL_HIP(hipEventRecord(e1_, s));
  readCtl(s);
  float ms = 0.f;
  NBL_HIP(hipEventElapsedTime(&ms, e0_, e1_));
  lastMs_ = ms;
  lastV_ = vcount_;
  reps_.resize(size_t(nb_));
  int most = 0;
  bool allConv = true;
  for (int q = 0; q < nb_; ++q) {
    const SolveCtl& h = hctl_[q];
    reps_[size_t(q)] = nbl_poisson_report{h.it, h.resid, h.converged};
    most = std::max(most, h.it);
    allConv = allConv && h.converged;
  }
  if (allConv && most > 0) predicted_ = most;
  return reps_[0];
}

void PoissonDev::profile(int kernel, int reps, hipStream_t s, double* ms, double* bytes) {
  if (batched_) throw std::invalid_argument("profile: not on a batch");
  if (kernel < 0 || kernel > 5) throw std::invalid_argument("profile: kernel must be 0 .. 5");
  if (reps < 1) throw std::invalid_argument("profile: reps must be >= 1");
  if (kind_ != 0) throw std::invalid_argument("profile: mg-pcg solver only");
  if ((kernel == 2 || kernel == 3) && nlevels() < 2) throw std::invalid_argument("profile: single-level grid");
  if (kernel == 4 && !tiled_) throw std::invalid_argument("profile: kernel 4 needs the two-kernels-per-level cycle");
  const MgLevel& L = dl_[0];
  const double n = double(L.n);
  const double sz = mixed_ ? 4.0 : 8.0;
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  // a live solve leaves ctl->done set; clear it (and every pending mean) so the
  // kernels do their work
  launch_ctl_init(ctl, 0.0, s);
  NBL_HIP(hipMemsetAsync(lx0_[0].as<void>(), 0, L.n * size_t(sz), s));
  NBL_HIP(hipMemsetAsync(lx1_[0].as<void>(), 0, L.n * size_t(sz), s));
  NBL_HIP(hipMemsetAsync(r_.as<void>(), 0, L.n * 8, s));
  NBL_HIP(hipMemsetAsync(p_.as<void>(), 0, L.n * 8, s));
  if (kernel == 5) {
    NBL_HIP(hipMemsetAsync(p2_.as<void>(), 0, L.n * 8, s));
    NBL_HIP(hipMemsetAsync(z_.as<void>(), 0, L.n * 8, s));
  }
  if (mixed_) NBL_HIP(hipMemsetAsync(b0f_.as<void>(), 0, L.n * 4, s));
  if (kernel == 3) {  // a non-trivial coarse right-hand side (mean-free for singular systems)
    const int lc = nlevels() - 1;
    const size_t nc = size_t(dl_[lc].n);
    std::vector<double> hb(nc);
    for (size_t c = 0; c < nc; ++c) hb[c] = std::sin(0.37 * double(c) + 0.11) * std::cos(0.013 * double(c));
    double mean = 0.0;
    size_t cnt = 0;
    for (size_t c = 0; c < nc; ++c)
      if (hl_[lc].code[c] & kFluidBit) {
        mean += hb[c];
        ++cnt;
      }
    mean = cnt ? mean / double(cnt) : 0.0;
    for (size_t c = 0; c < nc; ++c) hb[c] = (hl_[lc].code[c] & kFluidBit) ? hb[c] - (project_ ? mean : 0.0) : 0.0;
    if (mixed_) {
      std::vector<float> hf(hb.begin(), hb.end());
      NBL_HIP(hipMemcpyAsync(lb_[lc].as<void>(), hf.data(), nc * 4, hipMemcpyHostToDevice, s));
    } else {
      NBL_HIP(hipMemcpyAsync(lb_[lc].as<void>(), hb.data(), nc * 8, hipMemcpyHostToDevice, s));
    }
    NBL_HIP(hipStreamSynchronize(s));
  }
  const RedBuf rb = red();
  auto one = [&](int q) {
    if (kernel == 0) {
      if (mixed_)
        launch_mg_sweep<float>(L, b0f_.as<float>(), MeanRef{}, (q & 1) ? lx1_[0].as<float>() : lx0_[0].as<float>(),
                               MeanRef{}, false, nullptr, MeanRef{}, nullptr,
                               (q & 1) ? lx0_[0].as<float>() : lx1_[0].as<float>(), false, rb, ctl, s);
      else
        launch_mg_sweep<double>(L, r_.as<double>(), MeanRef{}, (q & 1) ? lx1_[0].as<double>() : lx0_[0].as<double>(),
                                MeanRef{}, false, nullptr, MeanRef{}, nullptr,
                                (q & 1) ? lx0_[0].as<double>() : lx1_[0].as<double>(), false, rb, ctl, s);
    } else if (kernel == 1) {
      launch_fcg_apply(L, p_.as<double>(), Ap_.as<double>(), ctl, rb, s);
    } else if (kernel == 5) {  // beta from SolveCtl (0 after the reset), p_old and p_new alternate
      launch_fcg_apply_pu(L, z_.as<double>(), (q & 1) ? p2_.as<double>() : p_.as<double>(),
                          (q & 1) ? p_.as<double>() : p2_.as<double>(), Ap_.as<double>(), ctl, rb, s, InFold{});
    } else if (kernel == 4) {
      if (mixed_) {
        DownArgs<float> d0;
        d0.L = dl_[0];
        d0.Lc = dl_[1];
        d0.ctl = ctl;
        d0.b = b0f_.as<float>();
        d0.xa = lx0_[0].as<float>();
        if (bf16_) d0.xaB = xaB_.as<__hip_bfloat16>();
        d0.bc = lb_[1].as<float>();
        d0.rowMean = project_ ? slotRow(slotB(1)) : nullptr;
        launch_down<float>(DS_B, d0, s);
      } else {
        DownArgs<double> d0;
        d0.L = dl_[0];
        d0.Lc = dl_[1];
        d0.ctl = ctl;
        d0.b = r_.as<double>();
        d0.xa = lx0_[0].as<double>();
        d0.bc = lb_[1].as<double>();
        d0.rowMean = project_ ? slotRow(slotB(1)) : nullptr;
        launch_down<double>(DS_B, d0, s);
      }
    } else if (kernel == 3) {
      const int lc = nlevels() - 1;
      if (mixed_) {
        const float* xo = nullptr;
        MeanRef mo{};
        vcycle<float>(lc, lb_[lc].as<float>(), MeanRef{}, &xo, &mo, s);
      } else {
        const double* xo = nullptr;
        MeanRef mo{};
        vcycle<double>(lc, lb_[lc].as<double>(), MeanRef{}, &xo, &mo, s);
      }
    } else {
      // as the per-sweep cycle runs it: the residual is folded only for a
      // coarsest level solved by CG on a singular system
      const bool rp = project_ && nlevels() == 2 && !direct_;
      if (mixed_)
        launch_mg_resid_restrict<float>(L, dl_[1], b0f_.as<float>(), MeanRef{}, lx0_[0].as<float>(), MeanRef{},
                                        lb_[1].as<float>(), rp, rb, ctl, s);
      else
        launch_mg_resid_restrict<double>(L, dl_[1], r_.as<double>(), MeanRef{}, lx0_[0].as<double>(), MeanRef{},
                                         lb_[1].as<double>(), rp, rb, ctl, s);
    }
  };
  one(0);  // warm-up
  NBL_HIP(hipEventRecord(e0_, s));
  for (int q = 0; q < reps; ++q) one(q);
  NBL_HIP(hipEventRecord(e1_, s));
  NBL_HIP(hipEventSynchronize(e1_));
  float t = 0.f;
  NBL_HIP(hipEventElapsedTime(&t, e0_, e1_));
  *ms = double(t) / reps;
  // algorithmic bytes: every input and output vector once, plus the 2-byte code
  const double nc = double(dl_.back().n);
  if (kernel == 0) *bytes = n * (3.0 * sz + 2.0);
  else if (kernel == 1) *bytes = n * (8.0 + 8.0 + 2.0);
  else if (kernel == 5) *bytes = n * (4.0 * 8.0 + 2.0);
  else if (kernel == 3) *bytes = direct_ ? nc * nc * sz + nc * (2.0 * sz + 2.0) : nc * (2.0 * sz + 2.0);
  else if (kernel == 4) *bytes = n * (sz + 2.0 + (bf16_ ? 2.0 : sz)) + double(dl_[1].n) * (sz + 2.0);
  else *bytes = n * (2.0 * sz + 2.0) + double(dl_[1].n) * (sz + 2.0);
}

// ---------------------------------------------------------------------------
// MAC run
// ---------------------------------------------------------------------------
RunDev::RunDev(nbl_ctx* ctx, const HostGrid& g, const nbl_flow_cfg& cfg, const nbl_solver_opts& opts)
    : ctx_(ctx), g_(g), cfg_(cfg) {
  if (!(cfg.reynolds > 0)) throw std::invalid_argument("reynolds must be > 0");
  NBL_HIP(hipStreamCreate(&stream_));
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  sz_[0] = size_t(nx + 1) * ny * nz;
  sz_[1] = size_t(nx) * (ny + 1) * nz;
  sz_[2] = size_t(nx) * ny * (nz + 1);
  sz_[3] = size_t(nx) * ny * nz;
  solid_ = DevBuf(ctx, sz_[3]);
  NBL_HIP(hipMemcpy(solid_.as<void>(), g.solid.data(), sz_[3], hipMemcpyHostToDevice));
  bool anySolid = false;
  for (uint8_t s : g.solid) anySolid |= s != 0;
  // faces touching a solid cell, per component (ns_solver.cpp:277-308)
  std::vector<int> list;
  auto sol = [&](int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return false;
    return g.solid[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)] != 0;
  };
  if (anySolid) {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i)
          if (sol(i - 1, j, k) || sol(i, j, k)) list.push_back(int(i + (nx + 1) * (j + ny * k)));
    solidCounts_[0] = int(list.size());
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i)
          if (sol(i, j - 1, k) || sol(i, j, k)) list.push_back(int(i + nx * (j + (ny + 1) * k)));
    solidCounts_[1] = int(list.size()) - solidCounts_[0];
    if (nz > 1)
      for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i)
            if (sol(i, j, k - 1) || sol(i, j, k)) list.push_back(int(i + nx * (j + ny * k)));
    solidCounts_[2] = int(list.size()) - solidCounts_[0] - solidCounts_[1];
    solidList_ = DevBuf(ctx, list.size() * sizeof(int));
    NBL_HIP(hipMemcpy(solidList_.as<void>(), list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  DevGrid& d = args_.g;
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.dx = g.dx;
  d.dy = g.dy;
  d.dz = g.dz;
  d.ox = g.org[0];
  d.oy = g.org[1];
  d.oz = g.org[2];
  for (int f = 0; f < 6; ++f) {
    d.bc[f] = g.face[f].type;
    for (int a = 0; a < 3; ++a) d.wall[f][a] = g.face[f].wall_velocity[a];
    d.profile[f] = g.face[f].profile;
    d.inletSpeed[f] = g.face[f].inlet_speed;
    d.inletAxis[f] = g.face[f].inlet_axis;
  }
  d.solid = solid_.as<uint8_t>();
  d.hasSolid = anySolid ? 1 : 0;
  args_.nu = cfg.velocity_scale * cfg.ref_length / cfg.reynolds;
  for (int a = 0; a < 3; ++a) args_.bodyForce[a] = cfg.body_force[a];
  args_.convection = cfg.convection;
  args_.cfl = cfg.cfl;
  args_.velocityScale = cfg.velocity_scale;
  args_.density = cfg.density;
  args_.refLength = cfg.ref_length;
  args_.refArea = cfg.ref_area;
  args_.tolContinuity = cfg.tol_continuity;
  args_.tolMass = cfg.tol_mass;
  for (int q = 0; q < 3; ++q) {
    bufA_[q] = DevBuf(ctx, sz_[q] * 8);
    bufB_[q] = DevBuf(ctx, sz_[q] * 8);
    buf0_[q] = DevBuf(ctx, sz_[q] * 8);
    cur_[q] = bufA_[q].as<double>();
    alt_[q] = bufB_[q].as<double>();
    NBL_HIP(hipMemset(cur_[q], 0, sz_[q] * 8));
  }
  p_ = DevBuf(ctx, sz_[3] * 8);
  b_ = DevBuf(ctx, sz_[3] * 8);
  phi_ = DevBuf(ctx, sz_[3] * 8);
  NBL_HIP(hipMemset(p_.as<void>(), 0, sz_[3] * 8));
  sc_ = DevBuf(ctx, sizeof(StepCtl));
  rep_ = DevBuf(ctx, sizeof(DevReport));
  pad_ = padLayout(args_.g);
  for (int c = 0; c < 3; ++c) {
    padBuf_[c] = DevBuf(ctx, size_t(pad_.n[c]) * 8);
    pad_.p[c] = padBuf_[c].as<double>();
  }
  gridDev_ = DevBuf(ctx, sizeof(DevGrid));
  NBL_HIP(hipMemcpy(gridDev_.as<void>(), &args_.g, sizeof(DevGrid), hipMemcpyHostToDevice));
  redStride_ = ns_partials(args_.g);
  red_ = DevBuf(ctx, sizeof(double) * kMaxRedVals * size_t(redStride_));
  NBL_HIP(hipMemset(sc_.as<void>(), 0, sizeof(StepCtl)));
  NBL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hsc_), sizeof(StepCtl), hipHostMallocDefault));
  NBL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hrep_), sizeof(DevReport), hipHostMallocDefault));
  std::memset(hsc_, 0, sizeof(StepCtl));
  if (opts.rhs_planes < 0 || opts.rhs_planes > 16)
    throw std::invalid_argument("rhs_planes must be 0 (automatic) or 1..16");
  rhsPlanes_ = opts.rhs_planes;
  nbl_solver_opts o = opts;
  poisson_ = std::make_unique<PoissonDev>(ctx, g, 0, cfg.poisson_tol, cfg.poisson_max_iters, o);
}

RunDev::~RunDev() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  poisson_.reset();
  if (hsc_) (void)hipHostFree(hsc_);
  if (hrep_) (void)hipHostFree(hrep_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void RunDev::bc(bool outflow) {
  launch_bc_planes(args_, cur_[0], cur_[1], cur_[2], outflow, stream_);
  if (args_.g.hasSolid)
    launch_bc_solid(args_, cur_[0], cur_[1], cur_[2], solidList_.as<int>(), solidCounts_, stream_);
}

void RunDev::initialize() {  // ns_solver.cpp:95-112
  for (int q = 0; q < 3; ++q) NBL_HIP(hipMemsetAsync(cur_[q], 0, sz_[q] * 8, stream_));
  NBL_HIP(hipMemsetAsync(p_.as<void>(), 0, sz_[3] * 8, stream_));
  for (int f = 0; f < 6; ++f)
    if (g_.face[f].type == NBL_BC_INLET && g_.face[f].inlet_axis == 0) {
      launch_fill(cur_[0], (long long)sz_[0], g_.face[f].inlet_speed, stream_);
      break;
    }
  bc(true);
  launch_set_step(sc_.as<StepCtl>(), 0, 0.0, 0.0, stream_);
  NBL_HIP(hipStreamSynchronize(stream_));
  dtLast_ = 0.0;
}

void RunDev::setState(const double* u, size_t nu, const double* v, size_t nv, const double* w, size_t nw,
                      const double* p, size_t np, int step, double t, double dtLast) {
  if (nu != sz_[0] || nv != sz_[1] || nw != sz_[2] || np != sz_[3])
    throw std::invalid_argument("NSSolver::setState: size mismatch");
  if (!u || !v || !w || !p) throw std::invalid_argument("setState: NULL array");
  const double* src[3] = {u, v, w};
  for (int q = 0; q < 3; ++q)
    NBL_HIP(hipMemcpyAsync(cur_[q], src[q], sz_[q] * 8, hipMemcpyHostToDevice, stream_));
  NBL_HIP(hipMemcpyAsync(p_.as<void>(), p, sz_[3] * 8, hipMemcpyHostToDevice, stream_));
  launch_set_step(sc_.as<StepCtl>(), step, t, dtLast, stream_);
  NBL_HIP(hipStreamSynchronize(stream_));
  dtLast_ = dtLast;
}

void RunDev::getState(double* u, double* v, double* w, double* p, int* step, double* t, double* dtLast) {
  double* dst[3] = {u, v, w};
  for (int q = 0; q < 3; ++q)
    if (dst[q]) NBL_HIP(hipMemcpyAsync(dst[q], cur_[q], sz_[q] * 8, hipMemcpyDeviceToHost, stream_));
  if (p) NBL_HIP(hipMemcpyAsync(p, p_.as<void>(), sz_[3] * 8, hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipMemcpyAsync(hsc_, sc_.as<void>(), sizeof(StepCtl), hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
  if (step) *step = hsc_->step;
  if (t) *t = hsc_->t;
  if (dtLast) *dtLast = dtLast_;
}

double RunDev::computeDt() {
  const RedBuf rb{red_.as<double>(), redStride_};
  launch_cfl(args_, cur_[0], cur_[1], cur_[2], sc_.as<StepCtl>(), 0.0, rb, stream_);
  NBL_HIP(hipMemcpyAsync(hsc_, sc_.as<void>(), sizeof(StepCtl), hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
  return hsc_->dt;
}

void RunDev::project(int stageIdx, nbl_poisson_report& rep) {  // ns_solver.cpp:475-579
  bc(true);
  launch_divergence(args_, cur_[0], cur_[1], cur_[2], b_.as<double>(), stream_);
  if (hostFn_) {  // a solver outside the library: host round trip
    const size_t n = sz_[3];
    hostB_.resize(n);
    hostX_.assign(n, 0.0);
    NBL_HIP(hipMemcpyAsync(hostB_.data(), b_.as<void>(), n * 8, hipMemcpyDeviceToHost, stream_));
    NBL_HIP(hipStreamSynchronize(stream_));
    rep = nbl_poisson_report{0, 0.0, 0};
    if (hostFn_(hostUser_, hostB_.data(), hostX_.data(), n, &rep) != 0)
      throw std::runtime_error("host pressure solver failed");
    NBL_HIP(hipMemcpyAsync(phi_.as<void>(), hostX_.data(), n * 8, hipMemcpyHostToDevice, stream_));
    NBL_HIP(hipStreamSynchronize(stream_));
    solveMs_[stageIdx - 1] = 0.0;
    solveIt_[stageIdx - 1] = rep.iterations;
    solveV_[stageIdx - 1] = 0;
  } else {
    rep = poisson_->solve(b_.as<double>(), phi_.as<double>(), stream_);
    solveMs_[stageIdx - 1] = poisson_->lastMs();
    solveIt_[stageIdx - 1] = rep.iterations;
    solveV_[stageIdx - 1] = poisson_->lastVcycles();
  }
  launch_correct(args_, cur_[0], cur_[1], cur_[2], phi_.as<double>(), p_.as<double>(), sc_.as<StepCtl>(),
                 stream_);
}

nbl_step_report RunDev::step(double dt) {  // ns_solver.cpp:765-824
  const RedBuf rb{red_.as<double>(), redStride_};
  StepCtl* sc = sc_.as<StepCtl>();
  launch_cfl(args_, cur_[0], cur_[1], cur_[2], sc, dt > 0.0 ? dt : 0.0, rb, stream_);
  double* u0[3] = {buf0_[0].as<double>(), buf0_[1].as<double>(), buf0_[2].as<double>()};
  for (int q = 0; q < 3; ++q)
    NBL_HIP(hipMemcpyAsync(u0[q], cur_[q], sz_[q] * 8, hipMemcpyDeviceToDevice, stream_));
  bc(true);
  nbl_poisson_report pr[3];
  for (int stage = 1; stage <= 3; ++stage) {
    launch_rhs_rk(args_, gridDev_.as<DevGrid>(), pad_, u0[0], u0[1], u0[2], cur_[0], cur_[1], cur_[2], alt_[0],
                  alt_[1], alt_[2], stage, sc, stream_, rhsPlanes_);
    std::swap(cur_, alt_);
    project(stage, pr[stage - 1]);
  }
  launch_diagnostics(args_, cur_[0], cur_[1], cur_[2], u0[0], u0[1], u0[2], p_.as<double>(), sc,
                     rep_.as<DevReport>(), rb, stream_);
  NBL_HIP(hipMemcpyAsync(hrep_, rep_.as<void>(), sizeof(DevReport), hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
  nbl_step_report r{};
  r.step = hrep_->step;
  r.t = hrep_->t;
  r.dt = hrep_->dt;
  r.cfl = hrep_->cfl;
  r.momentum_residual = hrep_->mom;
  r.continuity_residual = hrep_->divl2;
  r.divergence_max = hrep_->divmax;
  r.mass_error = hrep_->mass;
  r.poisson_iters = pr[0].iterations + pr[1].iterations + pr[2].iterations;
  r.poisson_residual = std::max({pr[0].residual, pr[1].residual, pr[2].residual});
  r.cd = hrep_->cd;
  r.cl = hrep_->cl;
  r.cell_count = hrep_->cells;
  r.accepted = hrep_->accepted;
  dtLast_ = r.dt;
  return r;
}

void RunDev::rhs(double* ru, double* rv, double* rw) {
  // R of the current state after applyVelocityBC (ns_solver.cpp:775-776), on a copy
  for (int q = 0; q < 3; ++q)
    NBL_HIP(hipMemcpyAsync(alt_[q], cur_[q], sz_[q] * 8, hipMemcpyDeviceToDevice, stream_));
  launch_bc_planes(args_, alt_[0], alt_[1], alt_[2], true, stream_);
  if (args_.g.hasSolid)
    launch_bc_solid(args_, alt_[0], alt_[1], alt_[2], solidList_.as<int>(), solidCounts_, stream_);
  double* out[3] = {buf0_[0].as<double>(), buf0_[1].as<double>(), buf0_[2].as<double>()};
  launch_rhs_rk(args_, gridDev_.as<DevGrid>(), pad_, alt_[0], alt_[1], alt_[2], alt_[0], alt_[1], alt_[2], out[0],
                out[1], out[2], 0, sc_.as<StepCtl>(), stream_, rhsPlanes_);
  double* dst[3] = {ru, rv, rw};
  for (int q = 0; q < 3; ++q) NBL_HIP(hipMemcpyAsync(dst[q], out[q], sz_[q] * 8, hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
}

void RunDev::cellVelocity(int comp, double* out) {
  if (comp < 0 || comp > 2) throw std::invalid_argument("component must be 0, 1 or 2");
  launch_cell_velocity(args_, cur_[0], cur_[1], cur_[2], comp, b_.as<double>(), stream_);
  NBL_HIP(hipMemcpyAsync(out, b_.as<void>(), sz_[3] * 8, hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
}

}  // namespace nbl
