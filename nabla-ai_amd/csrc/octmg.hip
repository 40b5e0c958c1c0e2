// octmg.hip — multigrid hierarchy and V-cycle on octree leaves.
//
// The reference solves the octree pressure system with Jacobi-PCG only
// (core/src/flow/octree_poisson.cpp:107-197); its design records name octree
// levels as the intended multigrid levels (docs/adr/0007, 0008) without an
// implementation.  This file builds that hierarchy on the device and applies
// the same V(2,2) cycle as the uniform solver (core/src/flow/multigrid.cpp:
// 102-194: damped Jacobi, omega 0.8, two pre- and two post-sweeps, exact-enough
// coarsest solve), with the restriction and prolongation of a
// volume-integrated finite-volume operator.
//
// Device layout per level: Morton-sorted leaves (OctLevel), up[i] = parent of
// leaf i in the next coarser level, cstart[I] = first child of coarse cell I
// in the next finer level (children are contiguous, at most 8).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "dev.hpp"
#include "octmg.hpp"

namespace nbl {
namespace {

constexpr double kOmega = 0.8;        // multigrid.cpp:38
constexpr size_t kCoarsestMax = 2048;  // multigrid.cpp:40
constexpr int kMaxLevels = 16;
constexpr int kOmBlocks = 4096;        // partial slots of the reducing sweep
constexpr int kRestrictCoarse = 32;    // coarse cells per restriction block (<= 8 children each)
constexpr int kFallbackSweeps = 16;    // coarsest level when its inverse cannot be formed (even)

#define OGRID(c, n) for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < (n); c += (long long)gridDim.x * blockDim.x)

__device__ __forceinline__ uint64_t clipKey(uint64_t key, int lev, int maxL, int clip) {
  return lev > clip ? key & ~((uint64_t(1) << (3 * (maxL - clip))) - 1) : key;
}

__global__ void k_clip_head(long long n, int maxL, int clip, const uint64_t* __restrict__ keys,
                            const uint8_t* __restrict__ lev, int32_t* __restrict__ head) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = clipKey(keys[i], lev[i], maxL, clip);
  head[i] = (i == 0 || clipKey(keys[i - 1], lev[i - 1], maxL, clip) != k) ? 1 : 0;
}

__global__ void k_clip_scatter(long long n, long long nc, int maxL, int clip, const uint64_t* __restrict__ keys,
                               const uint8_t* __restrict__ lev, const int32_t* __restrict__ head,
                               const int32_t* __restrict__ pos, uint64_t* __restrict__ ckeys,
                               uint8_t* __restrict__ clev, int32_t* __restrict__ cstart, int32_t* __restrict__ up) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t I = pos[i] - 1;
  up[i] = I;
  if (head[i]) {
    ckeys[I] = clipKey(keys[i], lev[i], maxL, clip);
    clev[I] = uint8_t(min(int(lev[i]), clip));
    cstart[I] = int32_t(i);
  }
  if (i == 0) cstart[nc] = int32_t(n);
}

// A coarse cell is fluid iff any child is fluid (multigrid.cpp:243-244).
__global__ void k_clip_mask(long long nc, const int32_t* __restrict__ cstart, const uint8_t* __restrict__ fmask,
                            uint8_t* __restrict__ cmask) {
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nc) return;
  uint8_t m = 2;
  for (int32_t f = cstart[I]; f < cstart[I + 1]; ++f)
    if (fmask[f] == 0) m = 0;
  cmask[I] = m;
}

struct PlainRd {
  const double* x;
  __device__ double operator()(long long j) const { return x[j]; }
};
struct ProlongRd {  // x + P e: each leaf adds its parent's correction
  const double* x;
  const double* e;
  const int32_t* up;
  __device__ double operator()(long long j) const { return x[j] + e[up[j]]; }
};

enum : int { OM_PLAIN = 0, OM_PROLONG = 1 };

// x <- x + omega (b - A x) / diag on fluid leaves, 0 elsewhere
// (multigrid.cpp:102-114); FOLD: per-block partial fluid sums of the result.
template <int MODE, bool FOLD>
__global__ void __launch_bounds__(kBlock) k_om_sweep(OctView L, const double* __restrict__ b,
                                                     const double* __restrict__ x, const double* __restrict__ ec,
                                                     const int32_t* __restrict__ up, double* __restrict__ xo,
                                                     RedBuf rb, const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  double acc[1] = {0.0};
  OGRID(i, L.n) {
    double v = 0.0;
    if (L.mask[i] == 0) {
      double xi, Ax;
      if constexpr (MODE == OM_PROLONG) {
        const ProlongRd X{x, ec, up};
        xi = X(i);
        Ax = octRowA(L, i, X);
      } else {
        const PlainRd X{x};
        xi = X(i);
        Ax = octRowA(L, i, X);
      }
      v = xi + kOmega * (b[i] - Ax) / L.diag[i];
      if (FOLD) acc[0] += v;
    }
    xo[i] = v;
  }
  if constexpr (FOLD) blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// Residual b - A x of the fine leaves, summed per coarse cell in child order
// (a block owns kRestrictCoarse coarse cells and their <= 8*kRestrictCoarse
// contiguous children).  With xc: the coarse level's first sweep from a zero
// guess, omega*bc/diag_c (fluid cells).
__global__ void __launch_bounds__(kBlock) k_om_restrict(OctView F, OctView C, const int32_t* __restrict__ cstart,
                                                        const double* __restrict__ b, const double* __restrict__ x,
                                                        double* __restrict__ bc, double* __restrict__ xc,
                                                        const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  __shared__ double rs[kRestrictCoarse * 8];
  const long long I0 = (long long)blockIdx.x * kRestrictCoarse;
  const long long I1 = min(I0 + kRestrictCoarse, C.n);
  const int32_t f0 = cstart[I0], f1 = cstart[I1];
  const long long i = f0 + (long long)threadIdx.x;
  double r = 0.0;
  if (i < f1 && F.mask[i] == 0) r = b[i] - octRowA(F, i, PlainRd{x});
  rs[threadIdx.x] = r;
  __syncthreads();
  const long long I = I0 + threadIdx.x;
  if (threadIdx.x < kRestrictCoarse && I < I1) {
    double s = 0.0;
    for (int32_t f = cstart[I]; f < cstart[I + 1]; ++f) s += rs[f - f0];
    bc[I] = s;
    if (xc) xc[I] = C.mask[I] == 0 ? kOmega * s / C.diag[I] : 0.0;
  }
}

__global__ void __launch_bounds__(kBlock) k_om_fsum(OctView L, const double* __restrict__ x, RedBuf rb,
                                                    const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  double acc[1] = {0.0};
  OGRID(i, L.n) if (L.mask[i] == 0) acc[0] += x[i];
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

__global__ void k_mask_code(long long n, const uint8_t* __restrict__ mask, uint16_t* __restrict__ code) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) code[i] = mask[i] == 0 ? kFluidBit : 0;
}

int omGrid(long long n) { return gridFor(n, kBlock, kOmBlocks); }

}  // namespace

OctMg::OctMg(nbl_ctx* ctx, const OctreeDev& fine, hipStream_t s) : ctx_(ctx) {
  hipEvent_t e0, e1;
  NBL_HIP(hipEventCreate(&e0));
  NBL_HIP(hipEventCreate(&e1));
  NBL_HIP(hipEventRecord(e0, s));
  lv_.reserve(kMaxLevels);
  lv_.emplace_back();
  lv_[0].op = &fine.level0();
  buildCoarse(fine, s);
  for (auto& L : lv_) {
    const size_t n = L.op->n;
    L.xa = DevBuf(ctx, n * 8);
    L.xb = DevBuf(ctx, n * 8);
    if (&L != &lv_[0]) L.b = DevBuf(ctx, n * 8);
  }
  direct_ = lv_.back().op->n <= kCoarsestMax && buildDirect(s);
  NBL_HIP(hipEventRecord(e1, s));
  NBL_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  NBL_HIP(hipEventElapsedTime(&t, e0, e1));
  buildMs_ = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  parts_ = omGrid((long long)lv_[0].op->n);
}

void OctMg::buildCoarse(const OctreeDev& fine, hipStream_t s) {
  const int maxL = fine.maxLevel();
  // finest level present in level 0
  const OctLevel& L0 = fine.level0();
  std::vector<uint8_t> hl(L0.n);
  NBL_HIP(hipMemcpyAsync(hl.data(), L0.lev.as<void>(), L0.n, hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  int top = 0;
  for (uint8_t v : hl) top = std::max(top, int(v));
  while (lv_.back().op->n > kCoarsestMax && int(lv_.size()) < kMaxLevels && top > 0) {
    const int clip = top - 1;
    Lvl& F = lv_.back();
    const OctLevel& fo = *F.op;
    const long long n = (long long)fo.n;
    DevBuf head(ctx_, fo.n * 4), pos(ctx_, fo.n * 4);
    k_clip_head<<<gridFor(n), kBlock, 0, s>>>(n, maxL, clip, fo.keys.as<uint64_t>(), fo.lev.as<uint8_t>(),
                                              head.as<int32_t>());
    NBL_LAUNCH_CHECK();
    size_t tb = 0;
    NBL_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, head.as<int32_t>(), pos.as<int32_t>(), int(n), s));
    DevBuf tmp(ctx_, tb);
    NBL_HIP(hipcub::DeviceScan::InclusiveSum(tmp.as<void>(), tb, head.as<int32_t>(), pos.as<int32_t>(), int(n), s));
    int32_t nc = 0;
    NBL_HIP(hipMemcpyAsync(&nc, pos.as<int32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    if (nc <= 0 || nc >= n) break;  // nothing merges
    Lvl C;
    C.own.n = size_t(nc);
    C.own.keys = DevBuf(ctx_, size_t(nc) * 8);
    C.own.lev = DevBuf(ctx_, size_t(nc));
    C.own.mask = DevBuf(ctx_, size_t(nc));
    C.cstart = DevBuf(ctx_, size_t(nc + 1) * 4);
    F.up = DevBuf(ctx_, fo.n * 4);
    k_clip_scatter<<<gridFor(n), kBlock, 0, s>>>(n, nc, maxL, clip, fo.keys.as<uint64_t>(), fo.lev.as<uint8_t>(),
                                                 head.as<int32_t>(), pos.as<int32_t>(), C.own.keys.as<uint64_t>(),
                                                 C.own.lev.as<uint8_t>(), C.cstart.as<int32_t>(), F.up.as<int32_t>());
    k_clip_mask<<<gridFor(nc), kBlock, 0, s>>>(nc, C.cstart.as<int32_t>(), fo.mask.as<uint8_t>(),
                                               C.own.mask.as<uint8_t>());
    NBL_LAUNCH_CHECK();
    // the restriction kernel stages at most 8 children per coarse cell
    std::vector<int32_t> hs(size_t(nc) + 1);
    NBL_HIP(hipMemcpyAsync(hs.data(), C.cstart.as<void>(), hs.size() * 4, hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    for (size_t I = 0; I < size_t(nc); ++I)
      if (hs[I + 1] - hs[I] < 1 || hs[I + 1] - hs[I] > 8)
        throw std::logic_error("octree multigrid: a coarse cell has more than 8 children");
    buildOctOperator(ctx_, maxL, fine.origin(), fine.fine(), C.own, nullptr, nullptr, s);
    lv_.push_back(std::move(C));
    lv_.back().op = &lv_.back().own;
    top = clip;
  }
}

// Dense inverse of the coarsest level (as the uniform solver's direct coarse
// solve): fluid rows carry the CSR couplings, other rows are identity; the
// always-singular fluid block gets alpha*11^T and the inverse is projected
// onto mean-free vectors afterwards.
bool OctMg::buildDirect(hipStream_t s) {
  const OctLevel& L = *lv_.back().op;
  const int n = int(L.n);
  std::vector<int32_t> rp(size_t(n) + 1), cl(std::max<size_t>(1, L.nnz));
  std::vector<double> cf(std::max<size_t>(1, L.nnz)), dg(L.n);
  std::vector<uint8_t> mk(L.n);
  NBL_HIP(hipMemcpyAsync(rp.data(), L.rowPtr.as<void>(), rp.size() * 4, hipMemcpyDeviceToHost, s));
  NBL_HIP(hipMemcpyAsync(cl.data(), L.col.as<void>(), L.nnz * 4, hipMemcpyDeviceToHost, s));
  NBL_HIP(hipMemcpyAsync(cf.data(), L.coef.as<void>(), L.nnz * 8, hipMemcpyDeviceToHost, s));
  NBL_HIP(hipMemcpyAsync(dg.data(), L.diag.as<void>(), size_t(n) * 8, hipMemcpyDeviceToHost, s));
  NBL_HIP(hipMemcpyAsync(mk.data(), L.mask.as<void>(), size_t(n), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  const size_t w = size_t(2) * n;
  std::vector<double> G(size_t(n) * w, 0.0);
  double dmax = 0.0, dfl = 0.0;
  for (int i = 0; i < n; ++i) {
    double* row = G.data() + size_t(i) * w;
    row[n + i] = 1.0;
    row[i] = dg[size_t(i)];
    if (mk[size_t(i)] != 0) continue;
    for (int e = rp[size_t(i)]; e < rp[size_t(i) + 1]; ++e) row[cl[size_t(e)]] -= cf[size_t(e)];
    dmax = std::max(dmax, dg[size_t(i)]);
    dfl += dg[size_t(i)];
  }
  if (L.nf > 0) {
    const double alpha = dfl / double(L.nf) / double(L.nf);
    for (int i = 0; i < n; ++i) {
      if (mk[size_t(i)] != 0) continue;
      double* row = G.data() + size_t(i) * w;
      for (int j = 0; j < n; ++j)
        if (mk[size_t(j)] == 0) row[j] += alpha;
    }
  }
  DevBuf g(ctx_, G.size() * sizeof(double)), colk(ctx_, size_t(n) * 8), rowk(ctx_, w * 8), fail(ctx_, 256);
  NBL_HIP(hipMemcpyAsync(g.as<void>(), G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice, s));
  NBL_HIP(hipMemsetAsync(fail.as<void>(), 0, 256, s));
  const double tiny = 1e-9 * std::max(dmax, 1.0);
  for (int k = 0; k < n; ++k) launch_gj_step(g.as<double>(), n, k, colk.as<double>(), rowk.as<double>(), fail.as<int>(), tiny, s);
  int failed = 0;
  NBL_HIP(hipMemcpyAsync(&failed, fail.as<void>(), sizeof(int), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  if (failed) return false;
  code_ = DevBuf(ctx_, size_t(n) * 2);
  k_mask_code<<<gridFor(n), kBlock, 0, s>>>(n, L.mask.as<uint8_t>(), code_.as<uint16_t>());
  NBL_LAUNCH_CHECK();
  M_ = DevBuf(ctx_, size_t(n) * n * sizeof(double));
  Mf_ = DevBuf(ctx_, size_t(n) * n * sizeof(float));
  DevBuf scratch(ctx_, size_t(n + 1) * sizeof(double));
  launch_inv_store(g.as<double>(), n, code_.as<uint16_t>(), (long long)L.nf, L.nf > 0, scratch.as<double>(),
                   M_.as<double>(), Mf_.as<float>(), s);
  NBL_HIP(hipStreamSynchronize(s));
  return true;
}

// V-cycle on level l with right-hand side b; the level's first sweep from a
// zero guess is already in xa.  *out receives the level's result.
void OctMg::cycle(size_t l, const double* b, RedBuf rb, const SolveCtl* ctl, hipStream_t s, const double** out) {
  Lvl& L = lv_[l];
  const OctView V = octView(*L.op);
  const long long n = V.n;
  const bool fold = l == 0;
  const bool last = l + 1 == lv_.size();
  double* xa = L.xa.as<double>();
  double* xb = L.xb.as<double>();
  if (last) {
    if (direct_) {
      MgLevel ml{};
      ml.n = n;
      ml.code = code_.as<uint16_t>();
      ml.nf = (long long)L.op->nf;
      launch_mg_coarse_direct<double>(ml, M_.as<double>(), Mf_.as<float>(), b, nullptr, xb, ctl, s);
    } else {
      // damped Jacobi from the first sweep already in xa; an odd number of
      // further sweeps ends in xb
      for (int q = 0; q < kFallbackSweeps - 1; ++q) {
        const bool even = (q & 1) == 0;
        k_om_sweep<OM_PLAIN, false><<<omGrid(n), kBlock, 0, s>>>(V, b, even ? xa : xb, nullptr, nullptr,
                                                                 even ? xb : xa, rb, ctl);
      }
      NBL_LAUNCH_CHECK();
    }
    if (fold) k_om_fsum<<<omGrid(n), kBlock, 0, s>>>(V, xb, rb, ctl);
    NBL_LAUNCH_CHECK();
    *out = xb;
    return;
  }
  // pre-smoothing: first sweep (in xa) then one plain sweep
  k_om_sweep<OM_PLAIN, false><<<omGrid(n), kBlock, 0, s>>>(V, b, xa, nullptr, nullptr, xb, rb, ctl);
  Lvl& C = lv_[l + 1];
  const OctView VC = octView(*C.op);
  const int rgrid = int((VC.n + kRestrictCoarse - 1) / kRestrictCoarse);
  k_om_restrict<<<rgrid, kBlock, 0, s>>>(V, VC, C.cstart.as<int32_t>(), b, xb, C.b.as<double>(),
                                         C.xa.as<double>(), ctl);
  NBL_LAUNCH_CHECK();
  const double* ec = nullptr;
  cycle(l + 1, C.b.as<double>(), rb, ctl, s, &ec);
  // post-smoothing: prolongation fused into the first sweep
  k_om_sweep<OM_PROLONG, false><<<omGrid(n), kBlock, 0, s>>>(V, b, xb, ec, L.up.as<int32_t>(), xa, rb, ctl);
  if (fold)
    k_om_sweep<OM_PLAIN, true><<<omGrid(n), kBlock, 0, s>>>(V, b, xa, nullptr, nullptr, xb, rb, ctl);
  else
    k_om_sweep<OM_PLAIN, false><<<omGrid(n), kBlock, 0, s>>>(V, b, xa, nullptr, nullptr, xb, rb, ctl);
  NBL_LAUNCH_CHECK();
  *out = xb;
}

void OctMg::vcycle(const double* r, RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  cycle(0, r, rb, ctl, s, &res0_);
}

}  // namespace nbl
