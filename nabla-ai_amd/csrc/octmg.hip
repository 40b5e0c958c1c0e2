// octmg.hip — multigrid hierarchy and V-cycle on octree leaves.
//
// The reference solves the octree pressure system with Jacobi-PCG only
// (core/src/flow/octree_poisson.cpp:107-197); its design records name octree
// levels as the intended multigrid levels (docs/adr/0007, 0008) without an
// implementation.  This file builds that hierarchy on the device and applies
// the same V(2,2) cycle as the uniform solver (core/src/flow/multigrid.cpp:
// 102-194: damped Jacobi, omega 0.8, two pre- and two post-sweeps, exact
// coarsest solve), with the restriction and prolongation of a
// volume-integrated finite-volume operator.
//
// Cells of every level are boxes in finest-index units (Morton key of the
// anchor, log2 size per axis).  A coarsening step raises a target log size S
// per axis: the axes whose physical target size is at most 1.5x the smallest
// grow by one, so cells of an anisotropic box (the cube-wake domain's 3:1:1
// leaves) first merge along the strongly coupled axes.  Cells no larger than
// S on every axis group by their S-aligned box; a group that covers its box
// exactly (volume sum) becomes one coarse cell, every other cell stays.  The
// coarse couplings sum the fine face areas between two coarse cells and use
// area / centre distance, the reference's face coefficient on octree cells.
//
// Device layout per level: OctLevel (keys, lsz, mask, CSR with areas and
// axes, slot-major operator), up[i] = parent of cell i in the next coarser
// level, and on coarse levels cstart[I] .. cstart[I+1] indexing child[], the
// cells of the next finer level that form cell I (increasing index).
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_reduce.hpp>

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
constexpr int kFallbackSweeps = 16;    // coarsest level when its inverse cannot be formed (even)

#define OGRID(c, n) for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < (n); c += (long long)gridDim.x * blockDim.x)

template <class T>
struct PlainRd {
  const T* x;
  __device__ T operator()(long long j) const { return x[j]; }
};
struct Bf16Rd {  // FP32 value of an iterate stored as BF16
  const __hip_bfloat16* x;
  __device__ float operator()(long long j) const { return __bfloat162float(x[j]); }
};
template <class T, class XR>
struct ProlongRd {  // x + P e: each leaf adds its parent's correction
  XR x;
  const T* e;
  const int32_t* up;
  __device__ T operator()(long long j) const { return x(j) + e[up[j]]; }
};

__device__ __forceinline__ void storeAs(double* p, double v) { *p = v; }
__device__ __forceinline__ void storeAs(float* p, float v) { *p = v; }
__device__ __forceinline__ void storeAs(__hip_bfloat16* p, float v) { *p = __float2bfloat16(v); }

// x <- x + omega (b - A x) / diag on fluid leaves, 0 elsewhere
// (multigrid.cpp:102-114), in T arithmetic with x read through X and the
// result stored as O.  FOLD 1: per-block partial fluid sums of the result;
// FOLD 2 (the cycle's last sweep inside the flexible CG): partials of
// result.rn, result.r and the fluid sum, in that row order.
template <class T, class XR, class O, int FOLD>
__global__ void __launch_bounds__(kBlock) k_om_sweep(OctViewT<T> L, const T* __restrict__ b, XR X,
                                                     O* __restrict__ xo, RedBuf rb, const SolveCtl* ctl,
                                                     const double* __restrict__ rn, const double* __restrict__ ro) {
  static_assert(FOLD >= 0 && FOLD <= 2, "fold mode");
  if (ctl && ldA(&ctl->done)) return;
  constexpr int K = FOLD == 2 ? 3 : 1;
  double acc[K] = {};
  OGRID(i, L.n) {
    // evaluated on every leaf (non-fluid rows have no couplings and a unit
    // diagonal), so no load waits for the mask
    const bool fl = L.mask[i] == 0;
    const T xi = X(i);
    const T Ax = octRowA(L, i, X);
    const T bi = b[i], di = L.diag[i];
    double rv = 0.0, r0 = 0.0;
    if constexpr (FOLD == 2) {
      rv = rn[i];
      r0 = ro[i];
    }
    T v = T(0);
    if (fl) {
      v = xi + T(kOmega) * (bi - Ax) / di;
      if constexpr (FOLD == 2) {
        acc[0] += double(v) * rv;
        acc[1] += double(v) * r0;
      }
      if constexpr (FOLD >= 1) acc[K - 1] += double(v);
    }
    storeAs(xo + i, v);
  }
  if constexpr (FOLD >= 1) blockToPartials<K>(acc, rb.partials, rb.stride, blockIdx.x);
}

// Residuals b - A x of a coarse cell's children, summed; with xc, the coarse
// level's first sweep from a zero guess, omega*bc/diag_c (fluid cells).
// G lanes of one wave per coarse cell: lane k takes children k, k+G, ... (one
// residual row each, in parallel instead of one after another), then the G
// lane sums are added by a butterfly in a fixed order, so the result does not
// depend on timing.  G is a power of two dividing the wave; every lane of a
// group shares I, so a group returns or shuffles as a whole.
template <class T, class XR, int G>
__global__ void __launch_bounds__(kBlock) k_om_restrict(OctViewT<T> F, OctViewT<T> C,
                                                        const int32_t* __restrict__ cstart,
                                                        const int32_t* __restrict__ child, const T* __restrict__ b,
                                                        XR X, T* __restrict__ bc, T* __restrict__ xc,
                                                        const SolveCtl* ctl) {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "lane group must be a power of two within a wave");
  if (ctl && ldA(&ctl->done)) return;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long I = t / G;
  const int k = int(t % G);
  if (I >= C.n) return;
  T s = T(0);
  const int32_t e = cstart[I + 1];
  for (int32_t p = cstart[I] + k; p < e; p += G) {
    const int32_t i = child[p];
    const T r = b[i] - octRowA(F, i, X);  // valid on every row; loads do not wait for the mask
    if (F.mask[i] == 0) s += r;
  }
#pragma unroll
  for (int o = 1; o < G; o <<= 1) s += __shfl_xor(s, o, G);
  if (k != 0) return;
  bc[I] = s;
  if (xc) xc[I] = C.mask[I] == 0 ? T(kOmega) * s / C.diag[I] : T(0);
}

// Lanes per coarse cell of a restriction: the power of two nearest the mean
// child count (1 .. 8).
int restrictLanes(long long nFine, long long nCoarse) {
  const double m = nCoarse > 0 ? double(nFine) / double(nCoarse) : 1.0;
  return m < 1.5 ? 1 : m < 3.0 ? 2 : m < 6.0 ? 4 : 8;
}

template <class T, class XR>
void launchRestrict(const OctViewT<T>& F, const OctViewT<T>& C, const int32_t* cstart, const int32_t* child,
                    const T* b, XR X, T* bc, T* xc, const SolveCtl* ctl, hipStream_t s) {
  const int G = restrictLanes(F.n, C.n);
  const int g = gridFor(C.n * G);
  switch (G) {
    case 1: k_om_restrict<T, XR, 1><<<g, kBlock, 0, s>>>(F, C, cstart, child, b, X, bc, xc, ctl); break;
    case 2: k_om_restrict<T, XR, 2><<<g, kBlock, 0, s>>>(F, C, cstart, child, b, X, bc, xc, ctl); break;
    case 4: k_om_restrict<T, XR, 4><<<g, kBlock, 0, s>>>(F, C, cstart, child, b, X, bc, xc, ctl); break;
    default: k_om_restrict<T, XR, 8><<<g, kBlock, 0, s>>>(F, C, cstart, child, b, X, bc, xc, ctl); break;
  }
}

__global__ void k_to_f32(long long n, const double* __restrict__ a, float* __restrict__ o) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = float(a[i]);
}

// ---- coarsening ---------------------------------------------------------------
__device__ __forceinline__ int lszAxis(uint32_t w, int ax) { return int((w >> (8 * ax)) & 0xffu); }

// Cells no larger than S on every axis: key of their S-aligned box and their
// volume (2^(lx+ly+lz)); other cells: the sentinel key and volume 0.
__global__ void k_sc_target(long long n, int S0, int S1, int S2, const uint64_t* __restrict__ keys,
                            const uint32_t* __restrict__ lsz, uint64_t* __restrict__ tkey,
                            uint64_t* __restrict__ vol, int32_t* __restrict__ idx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t w = lsz[i];
  const int l0 = lszAxis(w, 0), l1 = lszAxis(w, 1), l2 = lszAxis(w, 2);
  idx[i] = int32_t(i);
  if (l0 > S0 || l1 > S1 || l2 > S2) {
    tkey[i] = ~0ull;
    vol[i] = 0;
    return;
  }
  const uint64_t k = keys[i];
  tkey[i] = mortonEnc3(mortonAxis(k, 0) & ~((1u << S0) - 1u), mortonAxis(k, 1) & ~((1u << S1) - 1u),
                       mortonAxis(k, 2) & ~((1u << S2) - 1u));
  vol[i] = 1ull << (l0 + l1 + l2);
}

// Per group (run of equal target keys in sorted order): complete iff it is
// not the sentinel, has more than one member and its volumes fill the box.
__global__ void k_sc_complete(long long runs, const uint64_t* __restrict__ ukey, const int32_t* __restrict__ len,
                              const int32_t* __restrict__ off, const int32_t* __restrict__ idxS,
                              const uint64_t* __restrict__ vol, uint64_t full, int32_t* __restrict__ complete) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= runs) return;
  uint64_t v = 0;
  for (int32_t p = off[r]; p < off[r] + len[r]; ++p) v += vol[idxS[p]];
  complete[r] = (ukey[r] != ~0ull && len[r] > 1 && v == full) ? 1 : 0;
}

// Coarse key and size of every cell: its group's box when complete, else its own.
__global__ void k_sc_assign(long long runs, const uint64_t* __restrict__ ukey, const int32_t* __restrict__ len,
                            const int32_t* __restrict__ off, const int32_t* __restrict__ idxS,
                            const int32_t* __restrict__ complete, uint32_t Spacked,
                            const uint64_t* __restrict__ keys, const uint32_t* __restrict__ lsz,
                            uint64_t* __restrict__ ckey, uint32_t* __restrict__ clsz) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= runs) return;
  for (int32_t p = off[r]; p < off[r] + len[r]; ++p) {
    const int32_t i = idxS[p];
    ckey[i] = complete[r] ? ukey[r] : keys[i];
    clsz[i] = complete[r] ? Spacked : lsz[i];
  }
}

__global__ void k_iota32(long long n, int32_t* __restrict__ a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = int32_t(i);
}

__global__ void k_sc_heads(long long n, const uint64_t* __restrict__ k, int32_t* __restrict__ head) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) head[p] = (p == 0 || k[p] != k[p - 1]) ? 1 : 0;
}

// Parents from the cells sorted by coarse key: up, children ranges, coarse
// keys and sizes.
__global__ void k_sc_parents(long long n, long long nc, const uint64_t* __restrict__ ckeyS,
                             const int32_t* __restrict__ child, const int32_t* __restrict__ pid,
                             const int32_t* __restrict__ head, const uint32_t* __restrict__ clsz,
                             int32_t* __restrict__ up, int32_t* __restrict__ cstart, uint64_t* __restrict__ keysC,
                             uint32_t* __restrict__ lszC) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t I = pid[p] - 1;
  const int32_t i = child[p];
  up[i] = I;
  if (head[p]) {
    cstart[I] = int32_t(p);
    keysC[I] = ckeyS[p];
    lszC[I] = clsz[i];
  }
  if (p == 0) cstart[nc] = int32_t(n);
}

// A coarse cell is fluid iff any child is fluid (multigrid.cpp:243-244).
__global__ void k_sc_mask(long long nc, const int32_t* __restrict__ cstart, const int32_t* __restrict__ child,
                          const uint8_t* __restrict__ fmask, uint8_t* __restrict__ cmask) {
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nc) return;
  uint8_t m = 2;
  for (int32_t p = cstart[I]; p < cstart[I + 1]; ++p)
    if (fmask[child[p]] == 0) m = 0;
  cmask[I] = m;
}

// Coarse pair key of every fine coupling (sentinel inside one coarse cell).
__global__ void k_sc_emit(long long n, const int32_t* __restrict__ rowPtr, const int32_t* __restrict__ col,
                          const int32_t* __restrict__ up, uint64_t* __restrict__ ekey, int32_t* __restrict__ eid) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t I = uint32_t(up[i]);
  for (int32_t e = rowPtr[i]; e < rowPtr[i + 1]; ++e) {
    const uint32_t J = uint32_t(up[col[e]]);
    ekey[e] = I != J ? (uint64_t(I) << 32) | J : ~0ull;
    eid[e] = e;
  }
}

__global__ void k_sc_rowcount(long long nu, const uint64_t* __restrict__ ukey, int32_t* __restrict__ cnt) {
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < nu) atomicAdd(&cnt[ukey[u] >> 32], 1);
}

struct BoxGeo {
  double org[3], fine[3];
};
__device__ __forceinline__ double boxCentre(const BoxGeo& G, uint64_t key, uint32_t lsz, int ax) {
  return G.org[ax] + (double(mortonAxis(key, ax)) + 0.5 * double(1ull << lszAxis(lsz, ax))) * G.fine[ax];
}

// Coarse coupling u: the fine face areas of its run summed in fine coupling
// order, the run's axis, and coefficient area / centre distance.
__global__ void k_sc_fill(long long nu, const uint64_t* __restrict__ ukey, const int32_t* __restrict__ len,
                          const int32_t* __restrict__ off, const int32_t* __restrict__ eidS,
                          const double* __restrict__ farea, const uint8_t* __restrict__ fax, BoxGeo G,
                          const uint64_t* __restrict__ keysC, const uint32_t* __restrict__ lszC,
                          int32_t* __restrict__ col, double* __restrict__ coef, double* __restrict__ area,
                          uint8_t* __restrict__ cax) {
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  double a = 0.0;
  for (int32_t p = off[u]; p < off[u] + len[u]; ++p) a += farea[eidS[p]];
  const int ax = fax[eidS[off[u]]];
  const uint32_t I = uint32_t(ukey[u] >> 32), J = uint32_t(ukey[u] & 0xffffffffu);
  const double dist = fabs(boxCentre(G, keysC[I], lszC[I], ax) - boxCentre(G, keysC[J], lszC[J], ax));
  col[u] = int32_t(J);
  area[u] = a;
  cax[u] = uint8_t(ax);
  coef[u] = dist > 0.0 ? a / dist : 0.0;
}

__global__ void k_sc_diag(long long nc, const uint8_t* __restrict__ mask, const int32_t* __restrict__ rowPtr,
                          const double* __restrict__ coef, double* __restrict__ diag) {
  const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nc) return;
  double d = 0.0;
  if (mask[I] == 0)
    for (int32_t u = rowPtr[I]; u < rowPtr[I + 1]; ++u) d += coef[u];
  diag[I] = d > 0.0 ? d : 1.0;
}

// Partial fluid sums of x; with rn (the flexible CG's dots) the rows of
// k_om_sweep<FOLD 2>: x.rn, x.r, the fluid sum.
template <class T>
__global__ void __launch_bounds__(kBlock) k_om_fsum(long long n, const uint8_t* __restrict__ mask,
                                                    const T* __restrict__ x, RedBuf rb, const SolveCtl* ctl,
                                                    const double* __restrict__ rn, const double* __restrict__ ro) {
  if (ctl && ldA(&ctl->done)) return;
  if (rn) {
    double acc[3] = {0.0, 0.0, 0.0};
    OGRID(i, n) if (mask[i] == 0) {
      const double v = double(x[i]);
      acc[0] += v * rn[i];
      acc[1] += v * ro[i];
      acc[2] += v;
    }
    blockToPartials<3>(acc, rb.partials, rb.stride, blockIdx.x);
    return;
  }
  double acc[1] = {0.0};
  OGRID(i, n) if (mask[i] == 0) acc[0] += double(x[i]);
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

namespace {
// rocPRIM helpers: query the temporary size, allocate, run.
template <class F>
void cub(nbl_ctx* ctx, F&& f) {
  size_t tb = 0;
  NBL_HIP(f(nullptr, tb));
  DevBuf tmp(ctx, std::max<size_t>(tb, 1));
  NBL_HIP(f(tmp.as<void>(), tb));
}
template <class T>
T readBack(const T* d, hipStream_t s) {
  T h{};
  NBL_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  return h;
}
}  // namespace

void OctMg::buildCoarse(const OctreeDev& fine, hipStream_t s) {
  const int maxL = fine.maxLevel();
  BoxGeo G;
  for (int a = 0; a < 3; ++a) {
    G.org[a] = fine.origin()[a];
    G.fine[a] = fine.fine()[a];
  }
  // target size of level 0: the finest leaves
  {
    const OctLevel& L0 = fine.level0();
    std::vector<uint32_t> h(L0.n);
    NBL_HIP(hipMemcpyAsync(h.data(), L0.lsz.as<void>(), L0.n * 4, hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    int m = 1 << 30;
    for (uint32_t w : h) m = std::min(m, int(w & 0xffu));
    for (int a = 0; a < 3; ++a) lv_[0].S[a] = m;
  }
  while (lv_.back().op->n > kCoarsestMax && int(lv_.size()) < kMaxLevels) {
    Lvl& F = lv_.back();
    const OctLevel& fo = *F.op;
    const long long n = (long long)fo.n;
    const int ni = int(n);
    DevBuf tkey(ctx_, fo.n * 8), vol(ctx_, fo.n * 8), idx(ctx_, fo.n * 4), tkeyS(ctx_, fo.n * 8), idxS(ctx_, fo.n * 4);
    DevBuf ukey(ctx_, fo.n * 8), len(ctx_, fo.n * 4), off(ctx_, fo.n * 4), nrun(ctx_, 4), cmpl(ctx_, fo.n * 4);
    DevBuf ncmpl(ctx_, 4), ckey(ctx_, fo.n * 8), clsz(ctx_, fo.n * 4);
    int T[3] = {F.S[0], F.S[1], F.S[2]};
    bool merged = false;
    while (!(T[0] >= maxL && T[1] >= maxL && T[2] >= maxL)) {
      // grow the target along the axes whose physical size is within 1.5x of the smallest
      double h[3], hmin = 1e300;
      for (int a = 0; a < 3; ++a) {
        h[a] = G.fine[a] * double(1ull << T[a]);
        hmin = std::min(hmin, h[a]);
      }
      for (int a = 0; a < 3; ++a)
        if (h[a] <= 1.5 * hmin && T[a] < maxL) ++T[a];
      k_sc_target<<<gridFor(n), kBlock, 0, s>>>(n, T[0], T[1], T[2], fo.keys.as<uint64_t>(), fo.lsz.as<uint32_t>(),
                                                tkey.as<uint64_t>(), vol.as<uint64_t>(), idx.as<int32_t>());
      NBL_LAUNCH_CHECK();
      cub(ctx_, [&](void* t, size_t& tb) {
        return rocprim::radix_sort_pairs(t, tb, tkey.as<uint64_t>(), tkeyS.as<uint64_t>(), idx.as<int32_t>(),
                                                  idxS.as<int32_t>(), ni, 0, 64, s);
      });
      cub(ctx_, [&](void* t, size_t& tb) {
        return rocprim::run_length_encode(t, tb, tkeyS.as<uint64_t>(), (unsigned)(ni), ukey.as<uint64_t>(), len.as<int32_t>(), nrun.as<int32_t>(), s);
      });
      const int runs = readBack(nrun.as<int32_t>(), s);
      cub(ctx_, [&](void* t, size_t& tb) {
        return rocprim::exclusive_scan(t, tb, len.as<int32_t>(), off.as<int32_t>(), int32_t(0), size_t(runs), rocprim::plus<int32_t>(), s);
      });
      k_sc_complete<<<gridFor(runs), kBlock, 0, s>>>(runs, ukey.as<uint64_t>(), len.as<int32_t>(), off.as<int32_t>(),
                                                     idxS.as<int32_t>(), vol.as<uint64_t>(),
                                                     1ull << (T[0] + T[1] + T[2]), cmpl.as<int32_t>());
      NBL_LAUNCH_CHECK();
      cub(ctx_, [&](void* t, size_t& tb) {
        return rocprim::reduce(t, tb, cmpl.as<int32_t>(), ncmpl.as<int32_t>(), int32_t(0), size_t(runs), rocprim::plus<int32_t>(), s);
      });
      if (readBack(ncmpl.as<int32_t>(), s) == 0) continue;  // nothing merges at this size: grow further
      const uint32_t Sp = uint32_t(T[0]) | (uint32_t(T[1]) << 8) | (uint32_t(T[2]) << 16);
      k_sc_assign<<<gridFor(runs), kBlock, 0, s>>>(runs, ukey.as<uint64_t>(), len.as<int32_t>(), off.as<int32_t>(),
                                                   idxS.as<int32_t>(), cmpl.as<int32_t>(), Sp, fo.keys.as<uint64_t>(),
                                                   fo.lsz.as<uint32_t>(), ckey.as<uint64_t>(), clsz.as<uint32_t>());
      NBL_LAUNCH_CHECK();
      merged = true;
      break;
    }
    if (!merged) break;
    // cells sorted by coarse key: children grouped by parent, increasing index
    Lvl C;
    C.child = DevBuf(ctx_, fo.n * 4);
    DevBuf ckeyS(ctx_, fo.n * 8), head(ctx_, fo.n * 4), pid(ctx_, fo.n * 4);
    k_iota32<<<gridFor(n), kBlock, 0, s>>>(n, idx.as<int32_t>());
    NBL_LAUNCH_CHECK();
    cub(ctx_, [&](void* t, size_t& tb) {
      return rocprim::radix_sort_pairs(t, tb, ckey.as<uint64_t>(), ckeyS.as<uint64_t>(), idx.as<int32_t>(),
                                                C.child.as<int32_t>(), ni, 0, 64, s);
    });
    k_sc_heads<<<gridFor(n), kBlock, 0, s>>>(n, ckeyS.as<uint64_t>(), head.as<int32_t>());
    NBL_LAUNCH_CHECK();
    cub(ctx_, [&](void* t, size_t& tb) {
      return rocprim::inclusive_scan(t, tb, head.as<int32_t>(), pid.as<int32_t>(), size_t(ni), rocprim::plus<int32_t>(), s);
    });
    const int nc = readBack(pid.as<int32_t>() + (n - 1), s);
    if (nc <= 0 || nc >= n) break;
    OctLevel& co = C.own;
    co.n = size_t(nc);
    co.keys = DevBuf(ctx_, co.n * 8);
    co.lsz = DevBuf(ctx_, co.n * 4);
    co.mask = DevBuf(ctx_, co.n);
    C.cstart = DevBuf(ctx_, (co.n + 1) * 4);
    F.up = DevBuf(ctx_, fo.n * 4);
    k_sc_parents<<<gridFor(n), kBlock, 0, s>>>(n, nc, ckeyS.as<uint64_t>(), C.child.as<int32_t>(), pid.as<int32_t>(),
                                               head.as<int32_t>(), clsz.as<uint32_t>(), F.up.as<int32_t>(),
                                               C.cstart.as<int32_t>(), co.keys.as<uint64_t>(), co.lsz.as<uint32_t>());
    k_sc_mask<<<gridFor(nc), kBlock, 0, s>>>(nc, C.cstart.as<int32_t>(), C.child.as<int32_t>(), fo.mask.as<uint8_t>(),
                                             co.mask.as<uint8_t>());
    NBL_LAUNCH_CHECK();
    // coarse couplings: fine face areas summed per coarse pair
    const long long ne = (long long)fo.nnz;
    const int nei = int(ne);
    int nu = 0;
    DevBuf ekey(ctx_, std::max<size_t>(1, fo.nnz) * 8), eid(ctx_, std::max<size_t>(1, fo.nnz) * 4);
    DevBuf ekeyS(ctx_, std::max<size_t>(1, fo.nnz) * 8), eidS(ctx_, std::max<size_t>(1, fo.nnz) * 4);
    DevBuf uk(ctx_, std::max<size_t>(1, fo.nnz) * 8), ulen(ctx_, std::max<size_t>(1, fo.nnz) * 4);
    DevBuf uoff(ctx_, std::max<size_t>(1, fo.nnz) * 4);
    if (ne > 0) {
      k_sc_emit<<<gridFor(n), kBlock, 0, s>>>(n, fo.rowPtr.as<int32_t>(), fo.col.as<int32_t>(), F.up.as<int32_t>(),
                                              ekey.as<uint64_t>(), eid.as<int32_t>());
      NBL_LAUNCH_CHECK();
      cub(ctx_, [&](void* t, size_t& tb) {
        return rocprim::radix_sort_pairs(t, tb, ekey.as<uint64_t>(), ekeyS.as<uint64_t>(), eid.as<int32_t>(),
                                                  eidS.as<int32_t>(), nei, 0, 64, s);
      });
      cub(ctx_, [&](void* t, size_t& tb) {
        return rocprim::run_length_encode(t, tb, ekeyS.as<uint64_t>(), (unsigned)(nei), uk.as<uint64_t>(), ulen.as<int32_t>(), nrun.as<int32_t>(), s);
      });
      nu = readBack(nrun.as<int32_t>(), s);
      if (nu > 0 && readBack(uk.as<uint64_t>() + (nu - 1), s) == ~0ull) --nu;  // the sentinel run sorts last
      if (nu > 0)
        cub(ctx_, [&](void* t, size_t& tb) {
          return rocprim::exclusive_scan(t, tb, ulen.as<int32_t>(), uoff.as<int32_t>(), int32_t(0), size_t(nu), rocprim::plus<int32_t>(), s);
        });
    }
    co.nnz = size_t(nu);
    DevBuf rc(ctx_, (co.n + 1) * 4);
    co.rowPtr = DevBuf(ctx_, (co.n + 1) * 4);
    NBL_HIP(hipMemsetAsync(rc.as<void>(), 0, (co.n + 1) * 4, s));
    if (nu > 0) k_sc_rowcount<<<gridFor(nu), kBlock, 0, s>>>(nu, uk.as<uint64_t>(), rc.as<int32_t>());
    NBL_LAUNCH_CHECK();
    cub(ctx_, [&](void* t, size_t& tb) {
      return rocprim::exclusive_scan(t, tb, rc.as<int32_t>(), co.rowPtr.as<int32_t>(), int32_t(0), size_t(nc + 1), rocprim::plus<int32_t>(), s);
    });
    co.col = DevBuf(ctx_, std::max<size_t>(1, co.nnz) * 4);
    co.coef = DevBuf(ctx_, std::max<size_t>(1, co.nnz) * 8);
    co.area = DevBuf(ctx_, std::max<size_t>(1, co.nnz) * 8);
    co.cax = DevBuf(ctx_, std::max<size_t>(1, co.nnz));
    co.diag = DevBuf(ctx_, co.n * 8);
    if (nu > 0)
      k_sc_fill<<<gridFor(nu), kBlock, 0, s>>>(nu, uk.as<uint64_t>(), ulen.as<int32_t>(), uoff.as<int32_t>(),
                                               eidS.as<int32_t>(), fo.area.as<double>(), fo.cax.as<uint8_t>(), G,
                                               co.keys.as<uint64_t>(), co.lsz.as<uint32_t>(), co.col.as<int32_t>(),
                                               co.coef.as<double>(), co.area.as<double>(), co.cax.as<uint8_t>());
    k_sc_diag<<<gridFor(nc), kBlock, 0, s>>>(nc, co.mask.as<uint8_t>(), co.rowPtr.as<int32_t>(), co.coef.as<double>(),
                                             co.diag.as<double>());
    NBL_LAUNCH_CHECK();
    {
      std::vector<uint8_t> hm(co.n);
      NBL_HIP(hipMemcpyAsync(hm.data(), co.mask.as<void>(), co.n, hipMemcpyDeviceToHost, s));
      NBL_HIP(hipStreamSynchronize(s));
      co.nf = 0;
      for (uint8_t v : hm) co.nf += v == 0;
    }
    buildOctSlots(ctx_, co, s);
    for (int a = 0; a < 3; ++a) C.S[a] = T[a];
    lv_.push_back(std::move(C));
    lv_.back().op = &lv_.back().own;
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

void OctMg::setPrecision(int prec, hipStream_t s) {
  if (prec < NBL_PREC_FP64 || prec > NBL_PREC_MIXED_BF16) throw std::invalid_argument("unknown precision");
  prec_ = prec;
  if (prec == NBL_PREC_FP64 || lv_[0].xaF.bytes() > 0) {
    if (prec == NBL_PREC_MIXED_BF16 && lv_.size() > 1 && lv_[0].xbB.bytes() == 0)
      lv_[0].xbB = DevBuf(ctx_, lv_[0].op->n * sizeof(__hip_bfloat16));
    return;
  }
  auto toF = [&](const DevBuf& d, size_t n) {
    DevBuf f(ctx_, std::max<size_t>(1, n) * sizeof(float));
    if (n > 0) k_to_f32<<<gridFor((long long)n), kBlock, 0, s>>>((long long)n, d.as<double>(), f.as<float>());
    NBL_LAUNCH_CHECK();
    return f;
  };
  for (Lvl& L : lv_) {
    const OctLevel& o = *L.op;
    if (o.ncls) L.ctabF = toF(o.ctab, size_t(o.ncls));
    else L.scoefF = toF(o.scoef, size_t(kOctSlots) * o.n);
    L.xcoefF = toF(o.xcoef, o.nx);
    L.diagF = toF(o.diag, o.n);
    L.bF = DevBuf(ctx_, o.n * sizeof(float));
    L.xaF = DevBuf(ctx_, o.n * sizeof(float));
    L.xbF = DevBuf(ctx_, o.n * sizeof(float));
  }
  if (prec == NBL_PREC_MIXED_BF16 && lv_.size() > 1)
    lv_[0].xbB = DevBuf(ctx_, lv_[0].op->n * sizeof(__hip_bfloat16));
  NBL_HIP(hipStreamSynchronize(s));
}

const void* OctMg::diag0() const {
  return fp32() ? lv_[0].diagF.as<void>() : lv_[0].op->diag.as<void>();
}

template <>
OctViewT<double> OctMg::view<double>(size_t l) const {
  return octView(*lv_[l].op);
}
template <>
OctViewT<float> OctMg::view<float>(size_t l) const {
  const Lvl& L = lv_[l];
  const OctLevel& o = *L.op;
  return OctViewT<float>{(long long)o.n,          o.mask.as<uint8_t>(),  o.cnt.as<uint8_t>(),
                         o.scol.as<int32_t>(),    L.scoefF.as<float>(),  o.xrow.as<int32_t>(),
                         o.xcol.as<int32_t>(),    L.xcoefF.as<float>(),  L.diagF.as<float>(),
                         o.ncls ? o.meta.as<uint2>() : nullptr,          L.ctabF.as<float>()};
}

// V-cycle on level l with right-hand side b; the level's first sweep from a
// zero guess is already in xa.  *out receives the level's result.  On level
// 0 of the BF16 form the pre-smoothed iterate (written by the second
// pre-sweep, read by the restriction and the prolongating post-sweep) is
// stored as BF16.
template <class T>
void OctMg::cycle(size_t l, const T* b, RedBuf rb, const SolveCtl* ctl, hipStream_t s, const T** out) {
  constexpr bool kF = sizeof(T) == sizeof(float);
  Lvl& L = lv_[l];
  const OctViewT<T> V = view<T>(l);
  const long long n = V.n;
  const bool fold = l == 0;
  const bool last = l + 1 == lv_.size();
  T* xa = kF ? L.xaF.as<T>() : L.xa.as<T>();
  T* xb = kF ? L.xbF.as<T>() : L.xb.as<T>();
  // g: the reducing sweep's launch (at most kOmBlocks partials); gs: sweeps
  // without partials, one leaf per thread, so no thread walks several leaves
  // one dependent chain of loads after another
  const int g = omGrid(n);
  const int gs = gridFor(n);
  if (last) {
    if (direct_) {
      MgLevel ml{};
      ml.n = n;
      ml.code = code_.as<uint16_t>();
      ml.nf = (long long)L.op->nf;
      launch_mg_coarse_direct<T>(ml, M_.as<double>(), Mf_.as<float>(), b, MeanRef{}, xb, ctl, s);
    } else {
      // damped Jacobi from the first sweep already in xa; an odd number of
      // further sweeps ends in xb
      for (int q = 0; q < kFallbackSweeps - 1; ++q) {
        const bool even = (q & 1) == 0;
        k_om_sweep<T, PlainRd<T>, T, 0><<<gs, kBlock, 0, s>>>(V, b, PlainRd<T>{even ? xa : xb}, even ? xb : xa, rb,
                                                              ctl, nullptr, nullptr);
      }
      NBL_LAUNCH_CHECK();
    }
    if (fold) k_om_fsum<T><<<g, kBlock, 0, s>>>(n, V.mask, xb, rb, ctl, dRn_, dR_);
    NBL_LAUNCH_CHECK();
    *out = xb;
    return;
  }
  Lvl& C = lv_[l + 1];
  const OctViewT<T> VC = view<T>(l + 1);
  T* bc = kF ? C.bF.as<T>() : C.b.as<T>();
  T* xc = kF ? C.xaF.as<T>() : C.xa.as<T>();
  const T* ec = nullptr;
  auto post = [&](auto xr) {
    using R = ProlongRd<T, decltype(xr)>;
    k_om_sweep<T, R, T, 0><<<gs, kBlock, 0, s>>>(V, b, R{xr, ec, L.up.as<int32_t>()}, xa, rb, ctl, nullptr, nullptr);
    if (fold && dRn_)
      k_om_sweep<T, PlainRd<T>, T, 2><<<g, kBlock, 0, s>>>(V, b, PlainRd<T>{xa}, xb, rb, ctl, dRn_, dR_);
    else if (fold)
      k_om_sweep<T, PlainRd<T>, T, 1><<<g, kBlock, 0, s>>>(V, b, PlainRd<T>{xa}, xb, rb, ctl, nullptr, nullptr);
    else
      k_om_sweep<T, PlainRd<T>, T, 0><<<gs, kBlock, 0, s>>>(V, b, PlainRd<T>{xa}, xb, rb, ctl, nullptr, nullptr);
    NBL_LAUNCH_CHECK();
  };
  if constexpr (kF) {
    if (l == 0 && prec_ == NBL_PREC_MIXED_BF16) {
      __hip_bfloat16* xbB = L.xbB.as<__hip_bfloat16>();
      k_om_sweep<T, PlainRd<T>, __hip_bfloat16, 0><<<gs, kBlock, 0, s>>>(V, b, PlainRd<T>{xa}, xbB, rb, ctl, nullptr,
                                                                         nullptr);
      launchRestrict<T>(V, VC, C.cstart.as<int32_t>(), C.child.as<int32_t>(), b, Bf16Rd{xbB}, bc, xc, ctl, s);
      NBL_LAUNCH_CHECK();
      cycle<T>(l + 1, bc, rb, ctl, s, &ec);
      post(Bf16Rd{xbB});
      *out = xb;
      return;
    }
  }
  // pre-smoothing: first sweep (in xa) then one plain sweep
  k_om_sweep<T, PlainRd<T>, T, 0><<<gs, kBlock, 0, s>>>(V, b, PlainRd<T>{xa}, xb, rb, ctl, nullptr, nullptr);
  launchRestrict<T>(V, VC, C.cstart.as<int32_t>(), C.child.as<int32_t>(), b, PlainRd<T>{xb}, bc, xc, ctl, s);
  NBL_LAUNCH_CHECK();
  cycle<T>(l + 1, bc, rb, ctl, s, &ec);
  // post-smoothing: prolongation fused into the first sweep
  post(PlainRd<T>{xb});
  *out = xb;
}

void OctMg::vcycleDots(const double* r, const double* rOld, RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  dRn_ = r;
  dR_ = rOld;
  vcycle(r, rb, ctl, s);
  dRn_ = dR_ = nullptr;
}

void OctMg::vcycle(const double* r, RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  if (fp32()) {
    const float* out = nullptr;
    cycle<float>(0, lv_[0].bF.as<float>(), rb, ctl, s, &out);
    res0_ = out;
  } else {
    const double* out = nullptr;
    cycle<double>(0, r, rb, ctl, s, &out);
    res0_ = out;
  }
}

}  // namespace nbl
