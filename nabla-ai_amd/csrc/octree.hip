// octree.hip — octree leaves on the GPU (reference core/src/mesh/octree.cpp,
// core/include/nabla/mesh/morton.hpp, core/src/flow/octree_poisson.cpp).
//
// Layout in HBM: leaves Morton-sorted (SoA: key u64, level u8, mask u8).  The
// reference finds neighbours through a hash map probed at every operator
// application (octree.cpp:229-277); here the same probes run once, as binary
// searches over the sorted keys, and produce
//   * a face table (count + up to 4 neighbours per leaf and direction, in the
//     reference's order) — exported for the connectivity parity check, and
//   * a CSR operator: per fluid leaf, its fluid neighbours in the reference's
//     accumulation order with the exact face coefficient area/dist computed
//     from the reference's centre/size formulas, plus the diagonal.
// OctreePoisson::applyA is then one gather pass over the CSR, bit-identical to
// the reference per element.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "octmg.hpp"
#include "octree.hpp"

namespace nbl {
namespace {

// 21 bits per axis: x at bit 3i, y at 3i+1, z at 3i+2 (morton.hpp:15-48).
__host__ __device__ __forceinline__ uint64_t spread3(uint32_t a) {
  uint64_t x = a & 0x1fffffull;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__host__ __device__ __forceinline__ uint32_t gather3(uint64_t x) {
  x &= 0x1249249249249249ull;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
  x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
  x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
  x = (x ^ (x >> 32)) & 0x1fffffull;
  return uint32_t(x);
}
__host__ __device__ __forceinline__ uint64_t menc(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

__device__ __forceinline__ long long findKey(const uint64_t* __restrict__ keys, long long n, uint64_t k) {
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && keys[lo] == k) ? lo : -1;
}

// Neighbours across face `dir` of leaf i (octree.cpp:229-277).
__global__ void k_faces(long long n, int maxL, const uint64_t* __restrict__ keys, const uint8_t* __restrict__ lev,
                        uint8_t* __restrict__ cnt, int32_t* __restrict__ nbr) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * n) return;
  const long long i = t / 6;
  const int dir = int(t % 6);
  const int axis = dir >> 1;
  const bool plus = dir & 1;
  const uint64_t key = keys[i];
  const int L = lev[i];
  const uint32_t a[3] = {gather3(key), gather3(key >> 1), gather3(key >> 2)};
  const uint32_t size = 1u << (maxL - L);
  int64_t q[3] = {a[0], a[1], a[2]};
  q[axis] += plus ? int64_t(size) : -int64_t(size);
  int32_t* out = nbr + 4 * t;
  out[0] = out[1] = out[2] = out[3] = -1;
  if (q[axis] < 0 || q[axis] >= (int64_t(1) << maxL)) {
    cnt[t] = 0;
    return;
  }
  const uint32_t u[3] = {uint32_t(q[0]), uint32_t(q[1]), uint32_t(q[2])};
  for (int l = L; l >= 0; --l) {
    const uint32_t m = ~((1u << (maxL - l)) - 1u);
    const long long hit = findKey(keys, n, menc(u[0] & m, u[1] & m, u[2] & m));
    if (hit >= 0 && lev[hit] == l) {
      out[0] = int32_t(hit);
      cnt[t] = 1;
      return;
    }
  }
  int c = 0;
  if (L + 1 <= maxL) {
    const uint32_t cs = 1u << (maxL - L - 1);
    const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
    const uint32_t base = plus ? u[axis] : u[axis] + size - cs;
    for (int k1 = 0; k1 < 2; ++k1)
      for (int k2 = 0; k2 < 2; ++k2) {
        uint32_t b[3];
        b[axis] = base;
        b[t1] = u[t1] + uint32_t(k1) * cs;
        b[t2] = u[t2] + uint32_t(k2) * cs;
        const long long hit = findKey(keys, n, menc(b[0], b[1], b[2]));
        if (hit >= 0 && lev[hit] == L + 1) out[c++] = int32_t(hit);
      }
  }
  cnt[t] = uint8_t(c);
}

// area of the finer cell / centre distance (octree_poisson.cpp:24-32)
__device__ __forceinline__ double faceCoeff(const Geo& G, uint64_t ki, int li, uint64_t kj, int lj, int axis) {
  const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
  const int lf = lj > li ? lj : li;
  const double area = sizeAxis(G, lf, t1) * sizeAxis(G, lf, t2);
  const double dist = fabs(centerAxis(G, ki, li, axis) - centerAxis(G, kj, lj, axis));
  return dist > 0.0 ? area / dist : 0.0;
}

__global__ void k_row_count(long long n, const uint8_t* __restrict__ mask, const uint8_t* __restrict__ cnt,
                            const int32_t* __restrict__ nbr, int32_t* __restrict__ rowCount) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  if (mask[i] == 0) {
    for (int d = 0; d < 6; ++d)
      for (int m = 0; m < cnt[6 * i + d]; ++m)
        if (mask[nbr[24 * i + 4 * d + m]] == 0) ++c;
  }
  rowCount[i] = c;
}

// CSR rows plus, per coupling, the shared face area (the finer cell's
// transverse area) and the axis, which the multigrid coarsening sums.
__global__ void k_row_fill(long long n, Geo G, const uint64_t* __restrict__ keys, const uint8_t* __restrict__ lev,
                           const uint8_t* __restrict__ mask, const uint8_t* __restrict__ cnt,
                           const int32_t* __restrict__ nbr, const int32_t* __restrict__ rowPtr,
                           int32_t* __restrict__ col, double* __restrict__ coef, double* __restrict__ area,
                           uint8_t* __restrict__ cax, double* __restrict__ diag) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mask[i] != 0) {
    diag[i] = 1.0;
    return;
  }
  int e = rowPtr[i];
  double d = 0.0;
  for (int dd = 0; dd < 6; ++dd)
    for (int m = 0; m < cnt[6 * i + dd]; ++m) {
      const int j = nbr[24 * i + 4 * dd + m];
      if (mask[j] != 0) continue;
      const int ax = dd >> 1, t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
      const int lf = lev[j] > lev[i] ? lev[j] : lev[i];
      const double c = faceCoeff(G, keys[i], lev[i], keys[j], lev[j], ax);
      col[e] = j;
      coef[e] = c;
      area[e] = sizeAxis(G, lf, t1) * sizeAxis(G, lf, t2);
      cax[e] = uint8_t(ax);
      ++e;
      d += c;
    }
  diag[i] = d > 0.0 ? d : 1.0;
}

__global__ void k_lsz_from_level(long long n, int maxL, const uint8_t* __restrict__ lev, uint32_t* __restrict__ lsz) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t l = uint32_t(maxL - lev[i]);
  lsz[i] = l | (l << 8) | (l << 16);
}

struct XRd {  // x, or x - m with a pending mean
  const double* __restrict__ x;
  double m;
  bool hasM;
  __device__ __forceinline__ double operator()(long long j) const { return hasM ? x[j] - m : x[j]; }
};

// (A x)_i = sum_e coef_e (x_i - x_col_e) on fluid rows, x_i elsewhere
// (octree_poisson.cpp:64-85); with a pending mean non-fluid rows give 0.
__device__ __forceinline__ double octRow(const OctView& L, long long i, const double* __restrict__ x, double m,
                                         bool hasM) {
  // the row is evaluated on every leaf (a non-fluid row has no couplings),
  // so its loads do not wait for the mask
  const bool fl = L.mask[i] == 0;
  const double a = octRowA(L, i, XRd{x, m, hasM});
  return fl ? a : (hasM ? 0.0 : x[i]);
}

__global__ void k_oct_apply(OctView L, const double* __restrict__ x, double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  out[i] = octRow(L, i, x, 0.0, false);
}

// Slot-major copy of the CSR rows (first kOctSlots couplings per row) and the
// overflow counts.
__global__ void k_slot_count(long long n, const int32_t* __restrict__ rowPtr, uint8_t* __restrict__ cnt,
                             int32_t* __restrict__ xcount) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = rowPtr[i + 1] - rowPtr[i];
  cnt[i] = uint8_t(c);
  xcount[i] = c > kOctSlots ? c - kOctSlots : 0;
}
__global__ void k_slot_fill(long long n, const int32_t* __restrict__ rowPtr, const int32_t* __restrict__ col,
                            const double* __restrict__ coef, const int32_t* __restrict__ xrow,
                            int32_t* __restrict__ scol, double* __restrict__ scoef, int32_t* __restrict__ xcol,
                            double* __restrict__ xcoef) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r0 = rowPtr[i], c = rowPtr[i + 1] - r0;
  for (int s = 0; s < kOctSlots; ++s) {
    const long long q = (long long)s * n + i;
    scol[q] = s < c ? col[r0 + s] : int32_t(i);
    scoef[q] = s < c ? coef[r0 + s] : 0.0;
  }
  for (int e = kOctSlots; e < c; ++e) {
    xcol[xrow[i] + e - kOctSlots] = col[r0 + e];
    xcoef[xrow[i] + e - kOctSlots] = coef[r0 + e];
  }
}

__global__ void k_coef_bits(long long n, const double* __restrict__ coef, uint64_t* __restrict__ bits) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) bits[e] = uint64_t(__double_as_longlong(coef[e]));
}
// Per leaf: the class of every slot's coefficient (binary search of its bit
// pattern in the ascending class table; unused slots hold 0.0, which the
// table contains) and the row's coupling count, packed into 8 bytes.
__global__ void k_slot_meta(long long n, int ncls, const uint64_t* __restrict__ tab, const uint8_t* __restrict__ cnt,
                            const double* __restrict__ scoef, uint2* __restrict__ meta) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t w[2] = {0u, 0u};
  for (int s = 0; s < kOctSlots; ++s) {
    const uint64_t key = uint64_t(__double_as_longlong(scoef[(long long)s * n + i]));
    int lo = 0, hi = ncls - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tab[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    w[s >> 2] |= uint32_t(lo) << (8 * (s & 3));
  }
  w[1] |= uint32_t(cnt[i]) << 16;
  meta[i] = make_uint2(w[0], w[1]);
}

__global__ void k_gather(long long n, const uint64_t* __restrict__ perm, const double* __restrict__ src,
                         double* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
__global__ void k_scatter(long long n, const uint64_t* __restrict__ perm, const double* __restrict__ src,
                          double* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}

__global__ void k_iota(long long n, uint64_t* __restrict__ a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = uint64_t(i);
}
template <class T>
__global__ void k_permute(long long n, const uint64_t* __restrict__ perm, const T* __restrict__ src, T* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// Input validation: duplicates, misaligned anchors, levels out of range.
__global__ void k_validate(long long n, int maxL, const uint64_t* __restrict__ keys, const uint8_t* __restrict__ lev,
                           unsigned* __restrict__ bad) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int L = lev[i];
  if (L > maxL) {
    atomicOr(bad, 1u);
    return;
  }
  const uint32_t sz = 1u << (maxL - L);
  const uint64_t k = keys[i];
  const uint32_t a[3] = {gather3(k), gather3(k >> 1), gather3(k >> 2)};
  for (int q = 0; q < 3; ++q)
    if ((a[q] & (sz - 1u)) || a[q] + uint64_t(sz) > (uint64_t(1) << maxL)) atomicOr(bad, 2u);
  if (k >> 63) atomicOr(bad, 2u);
  if (i > 0 && keys[i - 1] == k) atomicOr(bad, 4u);
}

// ---- Jacobi-PCG on the leaves (octree_poisson.cpp:107-197) ------------------
// Every reducing kernel writes per-block partials (grid-stride over at most
// kOctBlocks blocks) and is followed by the shared finalize kernel (k_fin).
enum : int { O_MEAN_RHS = 0, O_MEAN_R = 1, O_MEAN_Z = 2, O_MEAN_X = 3, O_INTEGRAL = 4 };
constexpr int kOctBlocks = 4096;

#define OGRID(c, n) for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < (n); c += (long long)gridDim.x * blockDim.x)

__global__ void __launch_bounds__(kBlock) k_o_init(long long n, const uint8_t* __restrict__ mask,
                                                   const double* __restrict__ b, double* __restrict__ rhs,
                                                   double* __restrict__ x, RedBuf rb) {
  double acc[1] = {0.0};
  OGRID(c, n) {
    const double v = mask[c] == 0 ? b[c] : 0.0;
    rhs[c] = v;
    x[c] = 0.0;
    acc[0] += v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// rhs <- rhs - mean (fluid), 0 elsewhere; r = rhs; z' = r/diag (fluid);
// partials of |rhs|^2 and sum(z').
__global__ void __launch_bounds__(kBlock) k_o_init2(long long n, long long nf, const uint8_t* __restrict__ mask,
                                                    const double* __restrict__ diag, double* __restrict__ rhs,
                                                    double* __restrict__ r, double* __restrict__ z,
                                                    const SolveCtl* ctl, RedBuf rb) {
  const double m = ldA(&ctl->slot[O_MEAN_RHS]);
  double acc[2] = {0.0, 0.0};
  OGRID(c, n) {
    const bool fl = mask[c] == 0;
    const double v = (fl && nf) ? rhs[c] - m : (nf ? 0.0 : rhs[c]);
    rhs[c] = v;
    r[c] = v;
    const double zv = fl ? v / diag[c] : 0.0;
    z[c] = zv;
    acc[0] += v * v;
    if (fl) acc[1] += zv;
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, blockIdx.x);
}

// z <- z' - mean (fluid), 0 elsewhere; partials of r.z (first: also p = z).
__global__ void __launch_bounds__(kBlock) k_o_zdot(long long n, long long nf, const uint8_t* __restrict__ mask,
                                                   const double* __restrict__ r, double* __restrict__ z,
                                                   double* __restrict__ p, int first, const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double m = ldA(&ctl->slot[O_MEAN_Z]);
  double acc[1] = {0.0};
  OGRID(c, n) {
    const double zv = mask[c] == 0 ? (nf ? z[c] - m : z[c]) : 0.0;
    z[c] = zv;
    if (first) p[c] = zv;
    acc[0] += r[c] * zv;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock) k_o_apply_dot(OctView L, const double* __restrict__ p,
                                                        double* __restrict__ Ap, const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  double acc[1] = {0.0};
  OGRID(c, L.n) {
    const double v = octRow(L, c, p, 0.0, false);
    Ap[c] = v;
    acc[0] += p[c] * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// x += alpha p; r' = r - alpha Ap; partial fluid sums of r'.
__global__ void __launch_bounds__(kBlock) k_o_update(long long n, const uint8_t* __restrict__ mask,
                                                     double* __restrict__ x, const double* __restrict__ p,
                                                     double* __restrict__ r, const double* __restrict__ Ap,
                                                     const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double a = ldA(&ctl->alpha);
  double acc[1] = {0.0};
  OGRID(c, n) {
    x[c] = x[c] + a * p[c];
    const double v = r[c] - a * Ap[c];
    r[c] = v;
    if (mask[c] == 0) acc[0] += v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// r <- r' - mean (fluid), 0 elsewhere; z' = r/diag (fluid); partials of |r|^2
// and sum(z').
__global__ void __launch_bounds__(kBlock) k_o_rnorm(long long n, long long nf, const uint8_t* __restrict__ mask,
                                                    const double* __restrict__ diag, double* __restrict__ r,
                                                    double* __restrict__ z, const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double m = ldA(&ctl->slot[O_MEAN_R]);
  double acc[2] = {0.0, 0.0};
  OGRID(c, n) {
    const bool fl = mask[c] == 0;
    const double v = fl ? (nf ? r[c] - m : r[c]) : (nf ? 0.0 : r[c]);
    r[c] = v;
    const double zv = fl ? v / diag[c] : 0.0;
    z[c] = zv;
    acc[0] += v * v;
    if (fl) acc[1] += zv;
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock) k_o_pupdate(long long n, double* __restrict__ p, const double* __restrict__ z,
                                                      const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const double be = ldA(&ctl->beta);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) p[c] = z[c] + be * p[c];
}

__global__ void __launch_bounds__(kBlock) k_o_xsum(long long n, const uint8_t* __restrict__ mask,
                                                   const double* __restrict__ x, RedBuf rb) {
  double acc[1] = {0.0};
  OGRID(c, n) if (mask[c] == 0) acc[0] += x[c];
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock) k_o_final(OctView L, long long nf, const double* __restrict__ x,
                                                    const double* __restrict__ rhs, double* __restrict__ xo,
                                                    const SolveCtl* ctl, RedBuf rb) {
  const double m = ldA(&ctl->slot[O_MEAN_X]);
  const bool hasM = nf > 0;
  double acc[1] = {0.0};
  OGRID(c, L.n) {
    const double Ax = octRow(L, c, x, m, hasM);
    const double xv = L.mask[c] == 0 ? (hasM ? x[c] - m : x[c]) : (hasM ? 0.0 : x[c]);
    xo[c] = xv;
    const double e = rhs[c] - Ax;
    acc[0] += e * e;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock) k_count_fluid(long long n, const uint8_t* __restrict__ mask,
                                                        unsigned long long* __restrict__ cnt) {
  __shared__ unsigned sh[kBlock / 64];
  unsigned c = 0;
  OGRID(i, n) c += mask[i] == 0 ? 1u : 0u;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += sh[w];
    atomicAdd(cnt, t);
  }
}

// ---- multigrid-preconditioned flexible CG (multigrid.cpp:267-355 on leaves) --
constexpr double kOmegaFirst = 0.8;  // first V-cycle sweep from zero: omega*r/diag

// Level 0's right-hand side and first sweep of the V-cycle from residual
// value v of leaf c: FP64 cycle, omega*v/diag; FP32 cycle, its FP32 copy of v
// and omega*float(v)/diag in FP32 (the cycle's own arithmetic).
template <class T>
__device__ __forceinline__ void cycleInput(long long c, bool fl, double v, const double* diag, const T* diagT,
                                           T* rT, T* x0) {
  if constexpr (sizeof(T) == sizeof(float)) {
    const float vf = float(v);
    rT[c] = vf;
    x0[c] = fl ? float(kOmegaFirst) * vf / diagT[c] : 0.f;
  } else {
    x0[c] = fl ? kOmegaFirst * v / diag[c] : 0.0;
  }
}

// rhs <- rhs - mean (fluid), 0 elsewhere; r = rhs; the V-cycle's input;
// partials of |rhs|^2.
template <class T>
__global__ void __launch_bounds__(kBlock) k_om_init2(long long n, long long nf, const uint8_t* __restrict__ mask,
                                                     const double* __restrict__ diag, const T* __restrict__ diagT,
                                                     double* __restrict__ rhs, double* __restrict__ r,
                                                     T* __restrict__ rT, T* __restrict__ x0, const SolveCtl* ctl,
                                                     RedBuf rb) {
  const double m = ldA(&ctl->slot[O_MEAN_RHS]);
  double acc[1] = {0.0};
  OGRID(c, n) {
    const bool fl = mask[c] == 0;
    const double v = fl ? (nf ? rhs[c] - m : rhs[c]) : 0.0;
    rhs[c] = v;
    r[c] = v;
    cycleInput(c, fl, v, diag, diagT, rT, x0);
    acc[0] += v * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// Warm start, first pass: rhs = b and x = x0 on fluid leaves (0 elsewhere);
// partials of sum(rhs).
__global__ void __launch_bounds__(kBlock) k_o_init_warm(long long n, const uint8_t* __restrict__ mask,
                                                        const double* __restrict__ b, const double* __restrict__ x0,
                                                        double* __restrict__ rhs, double* __restrict__ x, RedBuf rb) {
  double acc[1] = {0.0};
  OGRID(c, n) {
    const bool fl = mask[c] == 0;
    const double v = fl ? b[c] : 0.0;
    rhs[c] = v;
    x[c] = fl ? x0[c] : 0.0;
    acc[0] += v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// Warm start, second pass: rhs <- rhs - mean (fluid), 0 elsewhere;
// r = rhs - A x; the V-cycle's input from r; partials of |rhs|^2 and |r|^2.
template <class T>
__global__ void __launch_bounds__(kBlock) k_om_init2_warm(OctView L, long long nf, const T* __restrict__ diagT,
                                                          const double* __restrict__ x, double* __restrict__ rhs,
                                                          double* __restrict__ r, T* __restrict__ rT,
                                                          T* __restrict__ x0, const SolveCtl* ctl, RedBuf rb) {
  const double m = ldA(&ctl->slot[O_MEAN_RHS]);
  double acc[2] = {0.0, 0.0};
  OGRID(c, L.n) {
    const bool fl = L.mask[c] == 0;
    const double v = fl ? (nf ? rhs[c] - m : rhs[c]) : 0.0;
    rhs[c] = v;
    const double rv = fl ? v - octRow(L, c, x, 0.0, false) : 0.0;
    r[c] = rv;
    cycleInput(c, fl, rv, L.diag, diagT, rT, x0);
    acc[0] += v * v;
    acc[1] += rv * rv;
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, blockIdx.x);
}

// z = V-cycle output - mean (fluid), 0 elsewhere; p = z; partials of r.z.
template <class T>
__global__ void __launch_bounds__(kBlock) k_om_start(long long n, long long nf, const uint8_t* __restrict__ mask,
                                                     const T* __restrict__ zc, const double* __restrict__ r,
                                                     double* __restrict__ z, double* __restrict__ p,
                                                     const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double m = ldA(&ctl->slot[O_MEAN_Z]);
  double acc[1] = {0.0};
  OGRID(c, n) {
    const double zo = double(zc[c]);
    const double zv = mask[c] == 0 ? (nf ? zo - m : zo) : 0.0;
    z[c] = zv;
    p[c] = zv;
    acc[0] += r[c] * zv;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// x += alpha p; rn = r - alpha Ap on fluid leaves, 0 elsewhere; the next
// V-cycle's input; partials of |rn|^2.  The right-hand side is projected once
// at the start and A annihilates constants on fluid leaves and is symmetric,
// so r - alpha Ap stays mean-free to rounding and rn needs no projection.
template <class T>
__global__ void __launch_bounds__(kBlock) k_om_update(long long n, const uint8_t* __restrict__ mask,
                                                      const double* __restrict__ diag, const T* __restrict__ diagT,
                                                      double* __restrict__ x, const double* __restrict__ p,
                                                      const double* __restrict__ r, const double* __restrict__ Ap,
                                                      double* __restrict__ rn, T* __restrict__ rT,
                                                      T* __restrict__ x0, const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double a = ldA(&ctl->alpha);
  double acc[1] = {0.0};
  OGRID(c, n) {
    const bool fl = mask[c] == 0;
    x[c] = x[c] + a * p[c];
    const double v = fl ? r[c] - a * Ap[c] : 0.0;
    rn[c] = v;
    cycleInput(c, fl, v, diag, diagT, rT, x0);
    acc[0] += v * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

// p = z + beta p with z = the V-cycle's output - its fluid mean (fluid
// leaves), 0 elsewhere: the projection of z formed where p reads it.
template <class T>
__global__ void __launch_bounds__(kBlock) k_om_pupdate(long long n, long long nf, const uint8_t* __restrict__ mask,
                                                       const T* __restrict__ zc, double* __restrict__ p,
                                                       const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const double be = ldA(&ctl->beta);
  const double m = ldA(&ctl->slot[O_MEAN_Z]);
  OGRID(c, n) {
    const double zo = double(zc[c]);
    const double zv = mask[c] == 0 ? (nf ? zo - m : zo) : 0.0;
    p[c] = zv + be * p[c];
  }
}

// ---- first-order conservative transport (adaptive/transport.cpp:24-60) ------
// One leaf per thread in sorted order, neighbours from the face table in the
// reference's order; every expression as the reference writes it.
__global__ void __launch_bounds__(kBlock) k_oct_transport(long long n, Geo G, const uint8_t* __restrict__ lev,
                                                          const uint8_t* __restrict__ cnt,
                                                          const int32_t* __restrict__ first,
                                                          const int32_t* __restrict__ xrow,
                                                          const int32_t* __restrict__ xnb,
                                                          const double* __restrict__ u, const double* __restrict__ v,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ phi, double dt,
                                                          double* __restrict__ phiOut) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int li = lev[i];
  const double szi[3] = {sizeAxis(G, li, 0), sizeAxis(G, li, 1), sizeAxis(G, li, 2)};
  const double vol = szi[0] * szi[1] * szi[2];
  const double pi = phi[i];
  int x = xrow[i];  // this leaf's further neighbours (directions with 2..4), in order
  double netOut = 0.0;
  for (int dir = 0; dir < 6; ++dir) {
    const int axis = dir >> 1;
    const double sign = (dir & 1) ? 1.0 : -1.0;  // +face vs -face normal
    const double* __restrict__ vel = axis == 0 ? u : axis == 1 ? v : w;
    const double vi = vel[i];
    const int c = cnt[6 * i + dir];
    for (int m = 0; m < c; ++m) {
      const int nb = m == 0 ? first[(long long)dir * n + i] : xnb[x++];
      const int ln = lev[nb];
      const double szn[3] = {sizeAxis(G, ln, 0), sizeAxis(G, ln, 1), sizeAxis(G, ln, 2)};
      const double areaI = axis == 0 ? szi[1] * szi[2] : axis == 1 ? szi[0] * szi[2] : szi[0] * szi[1];
      const double areaN = axis == 0 ? szn[1] * szn[2] : axis == 1 ? szn[0] * szn[2] : szn[0] * szn[1];
      const double area = areaN < areaI ? areaN : areaI;  // std::min
      const double vn = sign * 0.5 * (vi + vel[nb]);    // outward normal velocity
      const double up = vn >= 0.0 ? pi : phi[nb];
      netOut += vn * area * up;
    }
  }
  const double dphi = -dt / vol * netOut;
  phiOut[i] = pi + dphi;
}

// ---- adaptive sensors (PhysicalSensorEvaluator::evaluate, adaptive_core.cpp:73-176)
// One leaf per thread in sorted order.  The gradients across level jumps
// (Octree::gradient, octree.cpp:491-539) of u, v, w and p take one walk over the
// face table: per axis and side the neighbours' values and centres are summed
// in the reference's order, each field on its own, then averaged and
// differenced; the sensor formulas follow as the reference writes them.
// Results go to the caller's order through the permutation: out[f*n + perm[i]]
// for the eight real sensors and the carried speed / residual (rows 8, 9),
// flags[f*n + perm[i]] for near-wall, reverse flow and residual stagnation.

// Leaf centres (octree.cpp:84-91) per axis, cen[a*n + i], for the sensors'
// neighbour walks.
__global__ void k_oct_centers(long long n, Geo G, const uint64_t* __restrict__ keys, const uint8_t* __restrict__ lev,
                              double* __restrict__ cen) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; ++a) cen[a * n + i] = centerAxis(G, keys[i], lev[i], a);
}

__global__ void __launch_bounds__(kBlock) k_oct_sensors(long long n, Geo G, const double* __restrict__ cen,
                                                        const uint8_t* __restrict__ lev,
                                                        const uint8_t* __restrict__ mask, FaceWalk F,
                                                        const double* __restrict__ U, const double* __restrict__ V,
                                                        const double* __restrict__ W, const double* __restrict__ P,
                                                        const double* __restrict__ sPrev,
                                                        const double* __restrict__ rPrev, double nu,
                                                        const uint64_t* __restrict__ perm, double* __restrict__ out,
                                                        uint8_t* __restrict__ flags) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cnt[6], off[6];
  int x = F.xrow[i];
  for (int d = 0; d < 6; ++d) {
    cnt[d] = F.cnt[6 * i + d];
    off[d] = x;
    x += cnt[d] > 1 ? cnt[d] - 1 : 0;
  }
  auto nbAt = [&](int d, int m) { return m == 0 ? F.first[(long long)d * n + i] : F.xnb[off[d] + m - 1]; };
  const int li = lev[i];
  const double self[4] = {U[i], V[i], W[i], P[i]};
  double g[4][3];  // d(field)/d(axis) for u, v, w, p
  // One walk over the faces in direction order (-x, +x, -y, +y, -z, +z): the
  // gradient sums per side, and on the way the neighbours' speeds for the
  // truncation estimate (summed in the reference's direction order) and the
  // InsideSolid test of the near-wall flag.
  bool wall = cnt[2] == 0 || cnt[3] == 0;  // no leaf across a y face (floor, ceiling)
  double msum = 0.0;
  int mc = 0;
  for (int a = 0; a < 3; ++a) {
    const double* __restrict__ ca = cen + a * n;
    const double ci = ca[i];
    double val[2][4], pos[2] = {ci, ci};  // [0]: + face (2a+1), [1]: - face (2a)
    bool has[2] = {false, false};
    for (int t = 0; t < 2; ++t) {
      const int s = 1 - t;  // - face first
      const int d = 2 * a + t;
      for (int f = 0; f < 4; ++f) val[s][f] = self[f];
      const int c = cnt[d];
      if (c == 0) continue;
      double sv[4] = {0.0, 0.0, 0.0, 0.0}, sp = 0.0;
      for (int m = 0; m < c; ++m) {
        const int nb = nbAt(d, m);
        const double un = U[nb], vn = V[nb], wn = W[nb];
        sv[0] += un;
        sv[1] += vn;
        sv[2] += wn;
        sv[3] += P[nb];
        sp += ca[nb];
        if (mask[nb] == 2) wall = true;
        msum += sqrt(un * un + vn * vn + wn * wn);
        ++mc;
      }
      const double inv = 1.0 / double(c);
      for (int f = 0; f < 4; ++f) val[s][f] = sv[f] * inv;
      pos[s] = sp * inv;
      has[s] = true;
    }
    if (!has[0] && !has[1]) {
      for (int f = 0; f < 4; ++f) g[f][a] = 0.0;
      continue;
    }
    const double dd = pos[0] - pos[1];
    for (int f = 0; f < 4; ++f) g[f][a] = dd != 0.0 ? (val[0][f] - val[1][f]) / dd : 0.0;
  }
  const double* du = g[0];
  const double* dv = g[1];
  const double* dw = g[2];
  const double* dp = g[3];
  const double ox = dw[1] - dv[2], oy = du[2] - dw[0], oz = dv[0] - du[1];
  const double vort = sqrt(ox * ox + oy * oy + oz * oz);
  const double sxx = du[0], syy = dv[1], szz = dw[2];
  const double sxy = 0.5 * (du[1] + dv[0]), sxz = 0.5 * (du[2] + dw[0]), syz = 0.5 * (dv[2] + dw[1]);
  const double ss = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz);
  const double rxy = 0.5 * (du[1] - dv[0]), rxz = 0.5 * (du[2] - dw[0]), ryz = 0.5 * (dv[2] - dw[1]);
  const double oo = 2.0 * (rxy * rxy + rxz * rxz + ryz * ryz);
  const double gu = sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2] + dv[0] * dv[0] + dv[1] * dv[1] +
                         dv[2] * dv[2] + dw[0] * dw[0] + dw[1] * dw[1] + dw[2] * dw[2]);
  const double gp = sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
  const double speed = sqrt(self[0] * self[0] + self[1] * self[1] + self[2] * self[2]);
  double yplus = -1.0, tau = 0.0;
  if (wall) {
    const double yw = 0.5 * sizeAxis(G, li, 1);
    const double ut = sqrt(self[0] * self[0] + self[2] * self[2]);
    tau = nu * ut / (yw + 1e-300);
    yplus = sqrt(tau) * yw / (nu + 1e-300);
  }
  const double res = fabs(speed - sPrev[i]);
  const double rp = rPrev[i];
  const long long o = (long long)perm[i];
  out[o] = vort;
  out[n + o] = sqrt(2.0 * ss);
  out[2 * n + o] = 0.5 * (oo - ss);
  out[3 * n + o] = gp;
  out[4 * n + o] = gu;
  out[5 * n + o] = yplus;
  out[6 * n + o] = tau;
  out[7 * n + o] = mc ? fabs(speed - msum / mc) : 0.0;
  out[8 * n + o] = speed;
  out[9 * n + o] = res;
  flags[o] = wall ? 1 : 0;
  flags[n + o] = (wall && self[0] < 0.0) ? 1 : 0;
  flags[2 * n + o] = (rp > 1e-9 && res >= 0.95 * rp) ? 1 : 0;
}

// Face neighbours for the transport kernel: the first neighbour of every
// direction slot-major (first[dir*n + i], -1 when none), the rest of each
// direction (2..4 finer neighbours) per leaf in an overflow list.
__global__ void k_tab_count(long long n, const uint8_t* __restrict__ cnt, int32_t* __restrict__ extra) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int e = 0;
  for (int d = 0; d < 6; ++d) e += cnt[6 * i + d] > 1 ? cnt[6 * i + d] - 1 : 0;
  extra[i] = e;
}
__global__ void k_tab_fill(long long n, const uint8_t* __restrict__ cnt, const int32_t* __restrict__ nbr,
                           const int32_t* __restrict__ xrow, int32_t* __restrict__ first, int32_t* __restrict__ xnb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x = xrow[i];
  for (int d = 0; d < 6; ++d) {
    const int c = cnt[6 * i + d];
    first[(long long)d * n + i] = c > 0 ? nbr[24 * i + 4 * d] : -1;
    for (int m = 1; m < c; ++m) xnb[x++] = nbr[24 * i + 4 * d + m];
  }
}

// partials of sum_i phi_i vol_i (octreeFieldIntegral, transport.cpp:9-17)
__global__ void __launch_bounds__(kBlock) k_oct_integral(long long n, Geo G, const uint8_t* __restrict__ lev,
                                                         const double* __restrict__ phi, RedBuf rb) {
  double acc[1] = {0.0};
  OGRID(i, n) {
    const int l = lev[i];
    acc[0] += phi[i] * (sizeAxis(G, l, 0) * sizeAxis(G, l, 1) * sizeAxis(G, l, 2));
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}

int ogrid(long long n) { return gridFor(n, kBlock, kOctBlocks); }

}  // namespace

void buildOctSlots(nbl_ctx* ctx, OctLevel& L, hipStream_t s) {
  const size_t n = L.n;
  const long long N = (long long)n;
  L.cnt = DevBuf(ctx, n);
  L.scol = DevBuf(ctx, size_t(kOctSlots) * n * sizeof(int32_t));
  L.scoef = DevBuf(ctx, size_t(kOctSlots) * n * sizeof(double));
  L.xrow = DevBuf(ctx, (n + 1) * sizeof(int32_t));
  DevBuf xcount(ctx, (n + 1) * sizeof(int32_t));
  NBL_HIP(hipMemsetAsync(xcount.as<int32_t>() + n, 0, sizeof(int32_t), s));
  k_slot_count<<<gridFor(N), kBlock, 0, s>>>(N, L.rowPtr.as<int32_t>(), L.cnt.as<uint8_t>(), xcount.as<int32_t>());
  NBL_LAUNCH_CHECK();
  size_t tb = 0;
  NBL_HIP(rocprim::exclusive_scan(nullptr, tb, xcount.as<int32_t>(), L.xrow.as<int32_t>(), int32_t(0), size_t(int(n + 1)), rocprim::plus<int32_t>(), s));
  DevBuf tmp(ctx, tb);
  NBL_HIP(rocprim::exclusive_scan(tmp.as<void>(), tb, xcount.as<int32_t>(), L.xrow.as<int32_t>(), int32_t(0), size_t(int(n + 1)), rocprim::plus<int32_t>(), s));
  int32_t nxtra = 0;
  NBL_HIP(hipMemcpyAsync(&nxtra, L.xrow.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  L.nx = size_t(nxtra);
  L.xcol = DevBuf(ctx, std::max<size_t>(1, L.nx) * sizeof(int32_t));
  L.xcoef = DevBuf(ctx, std::max<size_t>(1, L.nx) * sizeof(double));
  k_slot_fill<<<gridFor(N), kBlock, 0, s>>>(N, L.rowPtr.as<int32_t>(), L.col.as<int32_t>(), L.coef.as<double>(),
                                            L.xrow.as<int32_t>(), L.scol.as<int32_t>(), L.scoef.as<double>(),
                                            L.xcol.as<int32_t>(), L.xcoef.as<double>());
  NBL_LAUNCH_CHECK();
  NBL_HIP(hipStreamSynchronize(s));
  buildOctClasses(ctx, L, s);
}

void buildOctClasses(nbl_ctx* ctx, OctLevel& L, hipStream_t s) {
  L.ncls = 0;
  L.meta = DevBuf();
  L.ctab = DevBuf();
  const size_t n = L.n, nnz = L.nnz;
  std::vector<uint64_t> tab{0ull};  // unused slots hold 0.0
  if (nnz > 0) {
    const long long E = (long long)nnz;
    DevBuf bits(ctx, nnz * sizeof(uint64_t)), sorted(ctx, nnz * sizeof(uint64_t)), uniq(ctx, nnz * sizeof(uint64_t));
    DevBuf nu(ctx, sizeof(unsigned));
    k_coef_bits<<<gridFor(E), kBlock, 0, s>>>(E, L.coef.as<double>(), bits.as<uint64_t>());
    NBL_LAUNCH_CHECK();
    size_t tb = 0, tb2 = 0;
    NBL_HIP(rocprim::radix_sort_keys(nullptr, tb, bits.as<uint64_t>(), sorted.as<uint64_t>(), nnz, 0, 64, s));
    NBL_HIP(rocprim::unique(nullptr, tb2, sorted.as<uint64_t>(), uniq.as<uint64_t>(), nu.as<unsigned>(), nnz,
                            rocprim::equal_to<uint64_t>(), s));
    DevBuf tmp(ctx, std::max(tb, tb2));
    NBL_HIP(rocprim::radix_sort_keys(tmp.as<void>(), tb, bits.as<uint64_t>(), sorted.as<uint64_t>(), nnz, 0, 64, s));
    NBL_HIP(rocprim::unique(tmp.as<void>(), tb2, sorted.as<uint64_t>(), uniq.as<uint64_t>(), nu.as<unsigned>(), nnz,
                            rocprim::equal_to<uint64_t>(), s));
    unsigned cnt = 0;
    NBL_HIP(hipMemcpyAsync(&cnt, nu.as<unsigned>(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    if (cnt > unsigned(kOctClasses)) return;
    std::vector<uint64_t> u(cnt);
    NBL_HIP(hipMemcpyAsync(u.data(), uniq.as<uint64_t>(), cnt * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    tab.insert(tab.end(), u.begin(), u.end());
    std::sort(tab.begin(), tab.end());
    tab.erase(std::unique(tab.begin(), tab.end()), tab.end());
    if (tab.size() > size_t(kOctClasses)) return;
  }
  L.ctab = DevBuf(ctx, tab.size() * sizeof(uint64_t));
  NBL_HIP(hipMemcpyAsync(L.ctab.as<void>(), tab.data(), tab.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  L.meta = DevBuf(ctx, std::max<size_t>(1, n) * sizeof(uint2));
  const long long N = (long long)n;
  k_slot_meta<<<gridFor(N), kBlock, 0, s>>>(N, int(tab.size()), L.ctab.as<uint64_t>(), L.cnt.as<uint8_t>(),
                                            L.scoef.as<double>(), L.meta.as<uint2>());
  NBL_LAUNCH_CHECK();
  NBL_HIP(hipStreamSynchronize(s));
  L.ncls = int(tab.size());
}

void buildOctOperator(nbl_ctx* ctx, int maxL, const double org[3], const double fine[3], OctLevel& L,
                      DevBuf* cntOut, DevBuf* nbrOut, hipStream_t s) {
  const size_t n = L.n;
  const long long N = (long long)n;
  DevBuf cnt(ctx, 6 * n), nbr(ctx, 24 * n * sizeof(int32_t));
  k_faces<<<gridFor(6 * N), kBlock, 0, s>>>(N, maxL, L.keys.as<uint64_t>(), L.lev.as<uint8_t>(), cnt.as<uint8_t>(),
                                            nbr.as<int32_t>());
  NBL_LAUNCH_CHECK();
  // CSR of the fluid-fluid couplings
  DevBuf rowCount(ctx, (n + 1) * sizeof(int32_t)), nfl(ctx, sizeof(unsigned long long));
  L.rowPtr = DevBuf(ctx, (n + 1) * sizeof(int32_t));
  NBL_HIP(hipMemsetAsync(rowCount.as<void>(), 0, (n + 1) * sizeof(int32_t), s));
  NBL_HIP(hipMemsetAsync(nfl.as<void>(), 0, sizeof(unsigned long long), s));
  k_row_count<<<gridFor(N), kBlock, 0, s>>>(N, L.mask.as<uint8_t>(), cnt.as<uint8_t>(), nbr.as<int32_t>(),
                                            rowCount.as<int32_t>());
  k_count_fluid<<<ogrid(N), kBlock, 0, s>>>(N, L.mask.as<uint8_t>(), nfl.as<unsigned long long>());
  NBL_LAUNCH_CHECK();
  size_t scanBytes = 0;
  NBL_HIP(rocprim::exclusive_scan(nullptr, scanBytes, rowCount.as<int32_t>(), L.rowPtr.as<int32_t>(), int32_t(0), size_t(int(n + 1)), rocprim::plus<int32_t>(), s));
  DevBuf scanTmp(ctx, scanBytes);
  NBL_HIP(rocprim::exclusive_scan(scanTmp.as<void>(), scanBytes, rowCount.as<int32_t>(), L.rowPtr.as<int32_t>(), int32_t(0), size_t(int(n + 1)), rocprim::plus<int32_t>(), s));
  int32_t nnz = 0;
  unsigned long long nf = 0;
  NBL_HIP(hipMemcpyAsync(&nnz, L.rowPtr.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipMemcpyAsync(&nf, nfl.as<void>(), sizeof(nf), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  L.nnz = size_t(nnz);
  L.nf = size_t(nf);
  L.col = DevBuf(ctx, std::max<size_t>(1, L.nnz) * sizeof(int32_t));
  L.coef = DevBuf(ctx, std::max<size_t>(1, L.nnz) * sizeof(double));
  L.diag = DevBuf(ctx, n * sizeof(double));
  Geo G;
  for (int a = 0; a < 3; ++a) {
    G.org[a] = org[a];
    G.fine[a] = fine[a];
  }
  G.maxL = maxL;
  L.area = DevBuf(ctx, std::max<size_t>(1, L.nnz) * sizeof(double));
  L.cax = DevBuf(ctx, std::max<size_t>(1, L.nnz));
  L.lsz = DevBuf(ctx, n * sizeof(uint32_t));
  k_row_fill<<<gridFor(N), kBlock, 0, s>>>(N, G, L.keys.as<uint64_t>(), L.lev.as<uint8_t>(), L.mask.as<uint8_t>(),
                                           cnt.as<uint8_t>(), nbr.as<int32_t>(), L.rowPtr.as<int32_t>(),
                                           L.col.as<int32_t>(), L.coef.as<double>(), L.area.as<double>(),
                                           L.cax.as<uint8_t>(), L.diag.as<double>());
  k_lsz_from_level<<<gridFor(N), kBlock, 0, s>>>(N, maxL, L.lev.as<uint8_t>(), L.lsz.as<uint32_t>());
  NBL_LAUNCH_CHECK();
  buildOctSlots(ctx, L, s);
  if (cntOut) *cntOut = std::move(cnt);
  if (nbrOut) *nbrOut = std::move(nbr);
  // a freed face table synchronises the device; the fill above has then completed
}

// ---------------------------------------------------------------------------
// OctreeDev
// ---------------------------------------------------------------------------
OctreeDev::OctreeDev(nbl_ctx* ctx, const uint64_t* keys, const uint8_t* levels, const uint8_t* masks, size_t n,
                     const double origin[3], const double extent[3], int maxLevel)
    : ctx_(ctx), maxL_(maxLevel) {
  if (n == 0) throw std::invalid_argument("octree: no leaves");
  if (maxLevel < 1 || maxLevel > 21) throw std::invalid_argument("Octree: maxLevel must be in [1, 21]");
  if (!(extent[0] > 0 && extent[1] > 0 && extent[2] > 0)) throw std::invalid_argument("Octree: extent must be positive");
  // The face table holds 24 entries per leaf and the coupling CSR up to 24 per
  // leaf, indexed in 32 bits: refuse trees whose 24 n would overflow them
  // before anything is allocated or launched.
  if (n > size_t(kMaxOctreeLeaves))
    throw CapacityError("octree: " + std::to_string(n) + " leaves exceed the 32-bit face-table limit of " +
                        std::to_string(kMaxOctreeLeaves));
  const double inv = 1.0 / double(1u << maxLevel);
  for (int a = 0; a < 3; ++a) {
    org_[a] = origin[a];
    fine_[a] = extent[a] * inv;
  }
  NBL_HIP(hipStreamCreate(&stream_));
  NBL_HIP(hipEventCreate(&e0_));
  NBL_HIP(hipEventCreate(&e1_));
  hipStream_t s = stream_;
  const long long N = (long long)n;
  DevBuf kin(ctx, n * 8), idx(ctx, n * 8), lin(ctx, n), min_(ctx, n);
  L0_.n = n;
  L0_.keys = DevBuf(ctx, n * 8);
  perm_ = DevBuf(ctx, n * 8);
  L0_.lev = DevBuf(ctx, n);
  L0_.mask = DevBuf(ctx, n);
  NBL_HIP(hipMemcpyAsync(kin.as<void>(), keys, n * 8, hipMemcpyHostToDevice, s));
  NBL_HIP(hipMemcpyAsync(lin.as<void>(), levels, n, hipMemcpyHostToDevice, s));
  NBL_HIP(hipMemcpyAsync(min_.as<void>(), masks, n, hipMemcpyHostToDevice, s));
  NBL_HIP(hipEventRecord(e0_, s));
  k_iota<<<gridFor(N), kBlock, 0, s>>>(N, idx.as<uint64_t>());
  NBL_LAUNCH_CHECK();
  size_t tmpBytes = 0;
  NBL_HIP(rocprim::radix_sort_pairs(nullptr, tmpBytes, kin.as<uint64_t>(), L0_.keys.as<uint64_t>(),
                                             idx.as<uint64_t>(), perm_.as<uint64_t>(), int(n), 0, 64, s));
  DevBuf tmp(ctx, tmpBytes);
  NBL_HIP(rocprim::radix_sort_pairs(tmp.as<void>(), tmpBytes, kin.as<uint64_t>(), L0_.keys.as<uint64_t>(),
                                             idx.as<uint64_t>(), perm_.as<uint64_t>(), int(n), 0, 64, s));
  k_permute<uint8_t><<<gridFor(N), kBlock, 0, s>>>(N, perm_.as<uint64_t>(), lin.as<uint8_t>(), L0_.lev.as<uint8_t>());
  k_permute<uint8_t><<<gridFor(N), kBlock, 0, s>>>(N, perm_.as<uint64_t>(), min_.as<uint8_t>(), L0_.mask.as<uint8_t>());
  NBL_LAUNCH_CHECK();
  DevBuf bad(ctx, 4);
  NBL_HIP(hipMemsetAsync(bad.as<void>(), 0, 4, s));
  k_validate<<<gridFor(N), kBlock, 0, s>>>(N, maxLevel, L0_.keys.as<uint64_t>(), L0_.lev.as<uint8_t>(), bad.as<unsigned>());
  NBL_LAUNCH_CHECK();
  unsigned hbad = 0;
  NBL_HIP(hipMemcpyAsync(&hbad, bad.as<void>(), 4, hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  if (hbad & 1u) throw std::invalid_argument("octree: level above maxLevel");
  if (hbad & 2u) throw std::invalid_argument("octree: key is not a valid anchor for its level");
  if (hbad & 4u) throw std::invalid_argument("octree: duplicate leaf key");
  buildOctOperator(ctx, maxLevel, org_, fine_, L0_, &cnt_, &nbr_, s);
  {
    DevBuf extra(ctx, (n + 1) * sizeof(int32_t));
    tfirst_ = DevBuf(ctx, 6 * n * sizeof(int32_t));
    txrow_ = DevBuf(ctx, (n + 1) * sizeof(int32_t));
    NBL_HIP(hipMemsetAsync(extra.as<int32_t>() + n, 0, sizeof(int32_t), s));
    k_tab_count<<<gridFor(N), kBlock, 0, s>>>(N, cnt_.as<uint8_t>(), extra.as<int32_t>());
    NBL_LAUNCH_CHECK();
    size_t tb = 0;
    NBL_HIP(rocprim::exclusive_scan(nullptr, tb, extra.as<int32_t>(), txrow_.as<int32_t>(), int32_t(0), size_t(int(n + 1)), rocprim::plus<int32_t>(), s));
    DevBuf tmp2(ctx, tb);
    NBL_HIP(rocprim::exclusive_scan(tmp2.as<void>(), tb, extra.as<int32_t>(), txrow_.as<int32_t>(), int32_t(0), size_t(int(n + 1)), rocprim::plus<int32_t>(), s));
    int32_t nx = 0;
    NBL_HIP(hipMemcpyAsync(&nx, txrow_.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    txnb_ = DevBuf(ctx, std::max<size_t>(1, size_t(nx)) * sizeof(int32_t));
    tnx_ = size_t(nx);
    k_tab_fill<<<gridFor(N), kBlock, 0, s>>>(N, cnt_.as<uint8_t>(), nbr_.as<int32_t>(), txrow_.as<int32_t>(),
                                             tfirst_.as<int32_t>(), txnb_.as<int32_t>());
    NBL_LAUNCH_CHECK();
  }
  NBL_HIP(hipEventRecord(e1_, s));
  // solver vectors
  for (auto* b : {&x_, &rhs_, &r_, &z_, &p_, &Ap_, &t0_, &t1_}) *b = DevBuf(ctx, n * 8);
  ctl_ = DevBuf(ctx, sizeof(SolveCtl));
  red_ = DevBuf(ctx, sizeof(double) * kMaxRedVals * kOctBlocks);
  NBL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hctl_), 2 * sizeof(SolveCtl), hipHostMallocDefault));
  NBL_HIP(hipStreamSynchronize(s));
  float t = 0.f;
  NBL_HIP(hipEventElapsedTime(&t, e0_, e1_));
  setupMs_ = t;
}

OctreeDev::~OctreeDev() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  mg_.reset();
  if (e0_) (void)hipEventDestroy(e0_);
  if (e1_) (void)hipEventDestroy(e1_);
  if (hctl_) (void)hipHostFree(hctl_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void OctreeDev::permutation(uint64_t* perm) {
  NBL_HIP(hipMemcpyAsync(perm, perm_.as<void>(), L0_.n * 8, hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
}

void OctreeDev::faceTable(uint8_t* count, int32_t* nbr) {
  if (count) NBL_HIP(hipMemcpyAsync(count, cnt_.as<void>(), 6 * L0_.n, hipMemcpyDeviceToHost, stream_));
  if (nbr) NBL_HIP(hipMemcpyAsync(nbr, nbr_.as<void>(), 24 * L0_.n * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
}

void OctreeDev::toSorted(const double* hostIn, double* dSorted) {
  NBL_HIP(hipMemcpyAsync(t1_.as<void>(), hostIn, L0_.n * 8, hipMemcpyHostToDevice, stream_));
  k_gather<<<gridFor((long long)L0_.n), kBlock, 0, stream_>>>((long long)L0_.n, perm_.as<uint64_t>(), t1_.as<double>(), dSorted);
  NBL_LAUNCH_CHECK();
}

void OctreeDev::fromSorted(const double* dSorted, double* hostOut) {
  k_scatter<<<gridFor((long long)L0_.n), kBlock, 0, stream_>>>((long long)L0_.n, perm_.as<uint64_t>(), dSorted, t1_.as<double>());
  NBL_LAUNCH_CHECK();
  NBL_HIP(hipMemcpyAsync(hostOut, t1_.as<void>(), L0_.n * 8, hipMemcpyDeviceToHost, stream_));
  NBL_HIP(hipStreamSynchronize(stream_));
}

void OctreeDev::applySorted(const double* x, double* out, hipStream_t s) {
  k_oct_apply<<<gridFor((long long)L0_.n), kBlock, 0, s>>>(octView(L0_), x, out);
  NBL_LAUNCH_CHECK();
}

void OctreeDev::apply(const double* x, double* out) {
  toSorted(x, t0_.as<double>());
  applySorted(t0_.as<double>(), Ap_.as<double>(), stream_);
  fromSorted(Ap_.as<double>(), out);
}

void OctreeDev::diagonal(double* out) { fromSorted(L0_.diag.as<double>(), out); }

double OctreeDev::transport(const double* u, const double* v, const double* w, double* phi, double dt, int steps) {
  if (steps < 1) throw std::invalid_argument("transport: steps must be >= 1");
  const long long N = (long long)L0_.n;
  hipStream_t s = stream_;
  toSorted(u, r_.as<double>());
  toSorted(v, z_.as<double>());
  toSorted(w, p_.as<double>());
  toSorted(phi, x_.as<double>());
  Geo G;
  for (int a = 0; a < 3; ++a) {
    G.org[a] = org_[a];
    G.fine[a] = fine_[a];
  }
  G.maxL = maxL_;
  double* a = x_.as<double>();
  double* b = Ap_.as<double>();
  NBL_HIP(hipEventRecord(e0_, s));
  for (int q = 0; q < steps; ++q) {
    k_oct_transport<<<gridFor(N), kBlock, 0, s>>>(N, G, L0_.lev.as<uint8_t>(), cnt_.as<uint8_t>(),
                                                  tfirst_.as<int32_t>(), txrow_.as<int32_t>(), txnb_.as<int32_t>(),
                                                  r_.as<double>(), z_.as<double>(), p_.as<double>(), a, dt, b);
    std::swap(a, b);
  }
  NBL_HIP(hipEventRecord(e1_, s));
  NBL_LAUNCH_CHECK();
  const RedBuf rb{red_.as<double>(), kOctBlocks};
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const int g = ogrid(N);
  k_oct_integral<<<g, kBlock, 0, s>>>(N, G, L0_.lev.as<uint8_t>(), a, rb);
  FinArgs fa{};
  fa.op = FIN_MEAN;
  fa.K = 1;
  fa.nparts = g;
  fa.slot = O_INTEGRAL;
  fa.nf = 1;  // the "mean" over one is the sum
  launch_fin(rb, fa, ctl, s);
  NBL_LAUNCH_CHECK();
  fromSorted(a, phi);  // synchronises the stream
  NBL_HIP(hipMemcpyAsync(hctl_, ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  float t = 0.f;
  NBL_HIP(hipEventElapsedTime(&t, e0_, e1_));
  lastMs_ = t;
  return hctl_[0].slot[O_INTEGRAL];
}

// Leaf centres for the sensors, computed on first use.
const double* OctreeDev::sensCen(hipStream_t s) {
  if (!cen_.bytes()) {
    const long long N = (long long)L0_.n;
    Geo G;
    for (int a = 0; a < 3; ++a) {
      G.org[a] = org_[a];
      G.fine[a] = fine_[a];
    }
    G.maxL = maxL_;
    cen_ = DevBuf(ctx_, 3 * L0_.n * sizeof(double));
    k_oct_centers<<<gridFor(N), kBlock, 0, s>>>(N, G, L0_.keys.as<uint64_t>(), L0_.lev.as<uint8_t>(),
                                                cen_.as<double>());
    NBL_LAUNCH_CHECK();
  }
  return cen_.as<double>();
}

void OctreeDev::sensors(const double* u, const double* v, const double* w, const double* p, double nu,
                        double* speedPrev, double* residPrev, double* out, uint8_t* flags) {
  const size_t n = L0_.n;
  const long long N = (long long)n;
  hipStream_t s = stream_;
  if (!sensOut_.bytes()) {
    sensOut_ = DevBuf(ctx_, 10 * n * sizeof(double));
    sensFlags_ = DevBuf(ctx_, 3 * n);
  }
  toSorted(u, r_.as<double>());
  toSorted(v, z_.as<double>());
  toSorted(w, p_.as<double>());
  toSorted(p, x_.as<double>());
  toSorted(speedPrev, rhs_.as<double>());
  toSorted(residPrev, Ap_.as<double>());
  Geo G;
  for (int a = 0; a < 3; ++a) {
    G.org[a] = org_[a];
    G.fine[a] = fine_[a];
  }
  G.maxL = maxL_;
  const FaceWalk F{cnt_.as<uint8_t>(), tfirst_.as<int32_t>(), txrow_.as<int32_t>(), txnb_.as<int32_t>()};
  NBL_HIP(hipEventRecord(e0_, s));
  k_oct_sensors<<<gridFor(N), kBlock, 0, s>>>(N, G, sensCen(s), L0_.lev.as<uint8_t>(),
                                              L0_.mask.as<uint8_t>(), F, r_.as<double>(), z_.as<double>(),
                                              p_.as<double>(), x_.as<double>(), rhs_.as<double>(), Ap_.as<double>(),
                                              nu, perm_.as<uint64_t>(), sensOut_.as<double>(),
                                              sensFlags_.as<uint8_t>());
  NBL_LAUNCH_CHECK();
  NBL_HIP(hipEventRecord(e1_, s));
  const double* d = sensOut_.as<double>();
  if (out) NBL_HIP(hipMemcpyAsync(out, d, 8 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipMemcpyAsync(speedPrev, d + 8 * n, n * sizeof(double), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipMemcpyAsync(residPrev, d + 9 * n, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (flags) NBL_HIP(hipMemcpyAsync(flags, sensFlags_.as<void>(), 3 * n, hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  float t = 0.f;
  NBL_HIP(hipEventElapsedTime(&t, e0_, e1_));
  lastMs_ = t;
}

void OctreeDev::profile(int kernel, int reps, double* ms, double* bytes) {
  if (kernel < 0 || kernel > 2)
    throw std::invalid_argument("octree profile: kernel must be 0 (SpMV), 1 (transport) or 2 (sensors)");
  if (reps < 1) throw std::invalid_argument("octree profile: reps must be >= 1");
  if (kernel == 2) {
    const size_t n = L0_.n;
    const long long N = (long long)n;
    if (!sensOut_.bytes()) {
      sensOut_ = DevBuf(ctx_, 10 * n * sizeof(double));
      sensFlags_ = DevBuf(ctx_, 3 * n);
    }
    for (DevBuf* d : {&r_, &z_, &p_, &x_, &rhs_, &Ap_}) NBL_HIP(hipMemsetAsync(d->as<void>(), 0, n * 8, stream_));
    Geo G;
    for (int a = 0; a < 3; ++a) {
      G.org[a] = org_[a];
      G.fine[a] = fine_[a];
    }
    G.maxL = maxL_;
    const FaceWalk F{cnt_.as<uint8_t>(), tfirst_.as<int32_t>(), txrow_.as<int32_t>(), txnb_.as<int32_t>()};
    const double* cen = sensCen(stream_);
    auto one = [&] {
      k_oct_sensors<<<gridFor(N), kBlock, 0, stream_>>>(
          N, G, cen, L0_.lev.as<uint8_t>(), L0_.mask.as<uint8_t>(), F, r_.as<double>(),
          z_.as<double>(), p_.as<double>(), x_.as<double>(), rhs_.as<double>(), Ap_.as<double>(), 1e-3,
          perm_.as<uint64_t>(), sensOut_.as<double>(), sensFlags_.as<uint8_t>());
    };
    one();
    NBL_HIP(hipEventRecord(e0_, stream_));
    for (int q = 0; q < reps; ++q) one();
    NBL_HIP(hipEventRecord(e1_, stream_));
    NBL_LAUNCH_CHECK();
    NBL_HIP(hipEventSynchronize(e1_));
    float t = 0.f;
    NBL_HIP(hipEventElapsedTime(&t, e0_, e1_));
    *ms = double(t) / reps;
    // compulsory traffic (a neighbour's centre, mask and fields are another
    // leaf's own): per leaf its three centre coordinates, level, mask, six
    // counts, six first-neighbour slots, u, v, w, p and the carried speed and
    // residual, the permutation, ten results and three flags; per further
    // neighbour its index
    *bytes = double(n) * (24.0 + 1.0 + 1.0 + 6.0 + 24.0 + 4.0 * 8.0 + 2.0 * 8.0 + 8.0 + 10.0 * 8.0 + 3.0) +
             double(tnx_) * 4.0;
    return;
  }
  if (kernel == 1) {
    const long long N = (long long)L0_.n;
    for (DevBuf* d : {&r_, &z_, &p_, &x_}) NBL_HIP(hipMemsetAsync(d->as<void>(), 0, L0_.n * 8, stream_));
    Geo G;
    for (int a = 0; a < 3; ++a) {
      G.org[a] = org_[a];
      G.fine[a] = fine_[a];
    }
    G.maxL = maxL_;
    auto one = [&](int q) {
      k_oct_transport<<<gridFor(N), kBlock, 0, stream_>>>(
          N, G, L0_.lev.as<uint8_t>(), cnt_.as<uint8_t>(), tfirst_.as<int32_t>(), txrow_.as<int32_t>(),
          txnb_.as<int32_t>(), r_.as<double>(), z_.as<double>(), p_.as<double>(),
          (q & 1) ? Ap_.as<double>() : x_.as<double>(), 0.01, (q & 1) ? x_.as<double>() : Ap_.as<double>());
    };
    one(0);
    NBL_HIP(hipEventRecord(e0_, stream_));
    for (int q = 0; q < reps; ++q) one(q + 1);
    NBL_HIP(hipEventRecord(e1_, stream_));
    NBL_LAUNCH_CHECK();
    NBL_HIP(hipEventSynchronize(e1_));
    float t = 0.f;
    NBL_HIP(hipEventElapsedTime(&t, e0_, e1_));
    *ms = double(t) / reps;
    // compulsory traffic (a neighbour's level, velocity and phi are another
    // leaf's own): per leaf its level, six counts, six first-neighbour slots,
    // the three velocities, phi in and out; per further neighbour its index
    *bytes = double(L0_.n) * (1.0 + 6.0 + 24.0 + 3.0 * 8.0 + 8.0 + 8.0) + double(tnx_) * 4.0;
    return;
  }
  NBL_HIP(hipMemsetAsync(t0_.as<void>(), 0, L0_.n * 8, stream_));
  applySorted(t0_.as<double>(), Ap_.as<double>(), stream_);
  NBL_HIP(hipEventRecord(e0_, stream_));
  for (int q = 0; q < reps; ++q) applySorted(t0_.as<double>(), Ap_.as<double>(), stream_);
  NBL_HIP(hipEventRecord(e1_, stream_));
  NBL_HIP(hipEventSynchronize(e1_));
  float t = 0.f;
  NBL_HIP(hipEventElapsedTime(&t, e0_, e1_));
  *ms = double(t) / reps;
  // coupling count, mask, x and y once per leaf; column index and coefficient
  // per coupling.  With coefficient classes: 8 bytes of classes and count,
  // mask, x and y per leaf, a column index per coupling, the FP64 coefficient
  // of each overflow coupling (beyond the six slots) and the class table.
  if (L0_.ncls)
    *bytes = double(L0_.n) * (8.0 + 1.0 + 8.0 + 8.0) + double(L0_.nnz) * 4.0 + double(L0_.nx) * 8.0 +
             double(L0_.ncls) * 8.0;
  else
    *bytes = double(L0_.n) * (1.0 + 1.0 + 8.0 + 8.0) + double(L0_.nnz) * (4.0 + 8.0);
}

nbl_poisson_report OctreeDev::solveSorted(const double* b, double* xo, double tol, int maxIters, hipStream_t s) {
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const RedBuf rb{red_.as<double>(), kOctBlocks};
  const long long N = (long long)L0_.n, nf = (long long)L0_.nf;
  const int g = ogrid(N);
  launch_ctl_init(ctl, tol, s);
  NBL_HIP(hipEventRecord(e0_, s));
  const uint8_t* mk = L0_.mask.as<uint8_t>();
  const double* dg = L0_.diag.as<double>();
  const OctView V = octView(L0_);
  double *x = x_.as<double>(), *rhs = rhs_.as<double>(), *r = r_.as<double>(), *z = z_.as<double>();
  double *p = p_.as<double>(), *Ap = Ap_.as<double>();
  auto fin = [&](int op, int K, int slot, bool pred) {
    FinArgs a{};
    a.op = op;
    a.K = K;
    a.nparts = g;
    a.slot = slot;
    a.nf = nf;
    a.convTol = 1e-6;
    a.predicated = pred ? 1 : 0;
    launch_fin(rb, a, ctl, s);
  };
  k_o_init<<<g, kBlock, 0, s>>>(N, mk, b, rhs, x, rb);
  fin(FIN_MEAN, 1, O_MEAN_RHS, false);
  k_o_init2<<<g, kBlock, 0, s>>>(N, nf, mk, dg, rhs, r, z, ctl, rb);
  fin(FIN_BNORM_MEAN, 2, O_MEAN_Z, false);
  k_o_zdot<<<g, kBlock, 0, s>>>(N, nf, mk, r, z, p, 1, ctl, rb);
  fin(FIN_RZ, 1, 0, true);
  NBL_LAUNCH_CHECK();
  int enq = 0;
  int batch = std::clamp(predicted_, 1, std::max(1, maxIters));
  while (enq < maxIters) {
    for (int q = 0; q < batch && enq < maxIters; ++q, ++enq) {
      k_o_apply_dot<<<g, kBlock, 0, s>>>(V, p, Ap, ctl, rb);
      fin(FIN_ALPHA, 1, 0, true);
      k_o_update<<<g, kBlock, 0, s>>>(N, mk, x, p, r, Ap, ctl, rb);
      fin(FIN_MEAN, 1, O_MEAN_R, true);
      k_o_rnorm<<<g, kBlock, 0, s>>>(N, nf, mk, dg, r, z, ctl, rb);
      fin(FIN_CONV_MEAN, 2, O_MEAN_Z, true);
      k_o_zdot<<<g, kBlock, 0, s>>>(N, nf, mk, r, z, p, 0, ctl, rb);
      fin(FIN_BETA, 1, 0, true);
      k_o_pupdate<<<gridFor(N), kBlock, 0, s>>>(N, p, z, ctl);
    }
    NBL_LAUNCH_CHECK();
    NBL_HIP(hipMemcpyAsync(hctl_, ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    if (hctl_[0].done) break;
    batch = std::max(8, batch / 4);
  }
  k_o_xsum<<<g, kBlock, 0, s>>>(N, mk, x, rb);
  fin(FIN_MEAN, 1, O_MEAN_X, false);
  k_o_final<<<g, kBlock, 0, s>>>(V, nf, x, rhs, xo, ctl, rb);
  fin(FIN_RESID, 1, 0, false);
  NBL_LAUNCH_CHECK();
  NBL_HIP(hipEventRecord(e1_, s));
  NBL_HIP(hipMemcpyAsync(hctl_, ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  float tms = 0.f;
  NBL_HIP(hipEventElapsedTime(&tms, e0_, e1_));
  lastMs_ = tms;
  const SolveCtl& h = hctl_[0];
  if (h.converged && h.it > 0) predicted_ = h.it;
  return nbl_poisson_report{h.it, h.resid, h.converged};
}

nbl_poisson_report OctreeDev::solve(const double* b, double* x, double tol, int maxIters) {
  if (maxIters < 0) throw std::invalid_argument("max iterations must be >= 0");
  // b is consumed by the first kernel of the solve, so the result may reuse its buffer
  toSorted(b, t0_.as<double>());
  const nbl_poisson_report rep = solveSorted(t0_.as<double>(), t0_.as<double>(), tol, maxIters, stream_);
  fromSorted(t0_.as<double>(), x);
  return rep;
}


OctMg& OctreeDev::mg() {
  if (!mg_) mg_ = std::make_unique<OctMg>(ctx_, *this, stream_);
  if (mg_->precision() != mgPrec_) mg_->setPrecision(mgPrec_, stream_);
  return *mg_;
}

void OctreeDev::setMgPrecision(int prec) {
  if (prec < NBL_PREC_FP64 || prec > NBL_PREC_MIXED_BF16)
    throw std::invalid_argument("precision must be NBL_PREC_FP64, NBL_PREC_MIXED or NBL_PREC_MIXED_BF16");
  mgPrec_ = prec;
}

int OctreeDev::mgLevels(long long* counts, int cap, double* buildMs) {
  OctMg& m = mg();
  for (int l = 0; l < std::min(cap, m.levels()); ++l) counts[l] = m.count(l);
  if (buildMs) *buildMs = m.buildMs();
  return m.levels();
}

nbl_poisson_report OctreeDev::solveMgSorted(const double* b, double* xo, double tol, int maxIters, hipStream_t s,
                                            const double* xs) {
  OctMg& M = mg();
  SolveCtl* ctl = ctl_.as<SolveCtl>();
  const RedBuf rb{red_.as<double>(), kOctBlocks};
  const long long N = (long long)L0_.n, nf = (long long)L0_.nf;
  const int g = ogrid(N);
  if (M.partials() != g) throw std::logic_error("octree multigrid: partial count mismatch");
  launch_ctl_init(ctl, tol, s);
  NBL_HIP(hipEventRecord(e0_, s));
  const uint8_t* mk = L0_.mask.as<uint8_t>();
  const double* dg = L0_.diag.as<double>();
  const OctView V = octView(L0_);
  double *x = x_.as<double>(), *rhs = rhs_.as<double>(), *r = r_.as<double>(), *z = z_.as<double>();
  double *p = p_.as<double>(), *Ap = Ap_.as<double>(), *rn = t1_.as<double>();
  auto fin = [&](int op, int K, int slot, bool pred, int early = 0) {
    FinArgs a{};
    a.op = op;
    a.K = K;
    a.nparts = g;
    a.slot = slot;
    a.nf = nf;
    a.convTol = tol;
    a.earlyExit = early;
    a.predicated = pred ? 1 : 0;
    launch_fin(rb, a, ctl, s);
  };
  if (xs) k_o_init_warm<<<g, kBlock, 0, s>>>(N, mk, b, xs, rhs, x, rb);
  else k_o_init<<<g, kBlock, 0, s>>>(N, mk, b, rhs, x, rb);
  fin(FIN_MEAN, 1, O_MEAN_RHS, false);
  // the V-cycle's level type: its input is written and its result read as double or float
  const bool f32 = M.fp32();
  float* rF = M.rhs0F();
  const float* dgF = static_cast<const float*>(M.diag0());
  if (xs) {
    if (f32)
      k_om_init2_warm<float><<<g, kBlock, 0, s>>>(V, nf, dgF, x, rhs, r, rF, static_cast<float*>(M.firstSweep()),
                                                  ctl, rb);
    else
      k_om_init2_warm<double><<<g, kBlock, 0, s>>>(V, nf, dg, x, rhs, r, nullptr,
                                                   static_cast<double*>(M.firstSweep()), ctl, rb);
    fin(FIN_BNORM_R0, 2, 0, false, 1);
  } else {
    if (f32)
      k_om_init2<float><<<g, kBlock, 0, s>>>(N, nf, mk, dg, dgF, rhs, r, rF, static_cast<float*>(M.firstSweep()),
                                             ctl, rb);
    else
      k_om_init2<double><<<g, kBlock, 0, s>>>(N, nf, mk, dg, dg, rhs, r, nullptr,
                                              static_cast<double*>(M.firstSweep()), ctl, rb);
    fin(FIN_BNORM, 1, 0, false, 1);
  }
  M.vcycle(r, rb, ctl, s);
  fin(FIN_MEAN, 1, O_MEAN_Z, true);
  if (f32)
    k_om_start<float><<<g, kBlock, 0, s>>>(N, nf, mk, static_cast<const float*>(M.result()), r, z, p, ctl, rb);
  else
    k_om_start<double><<<g, kBlock, 0, s>>>(N, nf, mk, static_cast<const double*>(M.result()), r, z, p, ctl, rb);
  fin(FIN_RZ, 1, 0, true);
  NBL_LAUNCH_CHECK();
  int enq = 0;
  int batch = std::clamp(predictedMg_, 1, std::max(1, maxIters));
  while (enq < maxIters) {
    // One flexible-CG iteration in seven launches: A p with p.Ap; alpha; x,
    // r_new and the cycle's input with |r_new|^2; the convergence test; the
    // V-cycle, whose last sweep folds z.r_new, z.r and z's fluid sum (r_new
    // and r are mean-free, so these are the projected z's dots); beta and z's
    // mean; p = (z - mean) + beta p.  r and r_new swap buffers.
    for (int q = 0; q < batch && enq < maxIters; ++q, ++enq) {
      k_o_apply_dot<<<g, kBlock, 0, s>>>(V, p, Ap, ctl, rb);
      fin(FIN_ALPHA, 1, 0, true);
      if (f32)
        k_om_update<float><<<g, kBlock, 0, s>>>(N, mk, dg, dgF, x, p, r, Ap, rn, rF,
                                                static_cast<float*>(M.firstSweep()), ctl, rb);
      else
        k_om_update<double><<<g, kBlock, 0, s>>>(N, mk, dg, dg, x, p, r, Ap, rn, nullptr,
                                                 static_cast<double*>(M.firstSweep()), ctl, rb);
      fin(FIN_CONV, 1, 0, true);
      M.vcycleDots(rn, r, rb, ctl, s);
      fin(FIN_BETA_FLEX_MEAN, 3, O_MEAN_Z, true);
      if (f32)
        k_om_pupdate<float><<<g, kBlock, 0, s>>>(N, nf, mk, static_cast<const float*>(M.result()), p, ctl);
      else
        k_om_pupdate<double><<<g, kBlock, 0, s>>>(N, nf, mk, static_cast<const double*>(M.result()), p, ctl);
      std::swap(r, rn);
    }
    NBL_LAUNCH_CHECK();
    NBL_HIP(hipMemcpyAsync(hctl_, ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    if (hctl_[0].done) break;
    batch = std::max(4, batch / 4);
  }
  k_o_xsum<<<g, kBlock, 0, s>>>(N, mk, x, rb);
  fin(FIN_MEAN, 1, O_MEAN_X, false);
  k_o_final<<<g, kBlock, 0, s>>>(V, nf, x, rhs, xo, ctl, rb);
  fin(FIN_RESID, 1, 0, false);
  NBL_LAUNCH_CHECK();
  NBL_HIP(hipEventRecord(e1_, s));
  NBL_HIP(hipMemcpyAsync(hctl_, ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s));
  NBL_HIP(hipStreamSynchronize(s));
  float tms = 0.f;
  NBL_HIP(hipEventElapsedTime(&tms, e0_, e1_));
  lastMs_ = tms;
  const SolveCtl& h = hctl_[0];
  if (h.converged && h.it > 0) predictedMg_ = h.it;
  return nbl_poisson_report{h.it, h.resid, h.converged};
}

nbl_poisson_report OctreeDev::solveMg(const double* b, double* x, double tol, int maxIters, bool warm) {
  if (maxIters < 0) throw std::invalid_argument("max iterations must be >= 0");
  mg();  // hierarchy set-up stays outside the timed solve
  toSorted(b, t0_.as<double>());
  // the initial guess waits in Ap_, which the iteration first writes after
  // the initial residual is formed
  if (warm) toSorted(x, Ap_.as<double>());
  const nbl_poisson_report rep = solveMgSorted(t0_.as<double>(), t0_.as<double>(), tol, maxIters, stream_,
                                               warm ? Ap_.as<double>() : nullptr);
  fromSorted(t0_.as<double>(), x);
  return rep;
}

}  // namespace nbl
