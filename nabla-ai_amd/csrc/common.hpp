// common.hpp — device/host plumbing shared by the gfx950 kernels.
//
// * HIP error handling that turns a failed runtime call into an nbl status and
//   a thread-local message (no exception crosses the C-ABI).
// * Deterministic grid reductions: every block folds its values with a fixed
//   wavefront/LDS tree and writes one partial per value; a one-block finalize
//   kernel then folds the partials in index order and applies the scalar
//   update (alpha, beta, a mean, the convergence test ...).  Launch shapes are
//   fixed per problem size, so results are bit-reproducible run to run.  No
//   atomics and no fences: a single ticket counter hit by thousands of blocks
//   serialises at the memory side and cost ~50 us per reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace nbl {

constexpr int kBlock = 256;          // threads per block for streaming kernels
constexpr int kWave = 64;            // CDNA wavefront
constexpr int kMaxRedVals = 8;       // values reduced together by one launch

// Work-group sizes read from the implicit kernel arguments (code object v5
// and v6, which the build pins:
// hidden_group_size_x, _y, _z are 16-bit fields at byte offsets 12, 14 and 16
// of the implicit argument block, which sits in device memory beside the
// kernel arguments): one scalar load.  In kernels that remap their block
// coordinates, blockDim.x can compile to a vector load at an offset selected
// by the block index (the remainder of a non-uniform grid), which every
// wavefront then waits for before its first address computation.  HIP launches
// uniform work-groups, so these are blockDim's values.  (The dispatch packet
// holds the same sizes, but it sits in the queue in host memory: reading it
// from every wavefront made the pressure solve 2.4x slower.)
__device__ __forceinline__ int wgSizeX() {
  return int(((const uint16_t*)__builtin_amdgcn_implicitarg_ptr())[6]);
}
__device__ __forceinline__ int wgSizeY() {
  return int(((const uint16_t*)__builtin_amdgcn_implicitarg_ptr())[7]);
}
__device__ __forceinline__ int wgSizeZ() {
  return int(((const uint16_t*)__builtin_amdgcn_implicitarg_ptr())[8]);
}

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// A request beyond a fixed capacity (an HBM budget, a 32-bit index range):
// NBL_ERR_CAPACITY at the C-ABI.
struct CapacityError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define NBL_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::nbl::HipError(std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

#define NBL_LAUNCH_CHECK() NBL_HIP(hipGetLastError())

inline int gridFor(long long n, int block = kBlock, int cap = 1 << 30) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

// Scratch for deterministic reductions: kMaxRedVals rows of `stride` partials
// (one per block of the largest reducing launch).
struct RedBuf {
  double* partials;
  int stride;
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
// Scalars that one kernel hands to the next (per-block partials, the solve's
// alpha/beta/means/flags) are written and read with agent-scope (device-wide
// coherent) accesses: the producer is a single block on one XCD and the
// consumers run on all eight, each XCD with its own L2.
template <class T>
__device__ __forceinline__ T ldA(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void stA(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wavefront reductions with DPP lane moves (no LDS traffic): pairs and quads
// (quad_perm), rows of 16 (row_ror 4, 8), then row_bcast15 / row_bcast31
// carry rows 0..2 into row 3, so lane 63 ends with the total, which readlane
// hands to every lane.  The order is fixed, so results are reproducible.
// Every lane of the wavefront must be active.
template <int CTRL, int ROWMASK, bool MAX>
__device__ __forceinline__ double dppCombine(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const int lo = int(unsigned(b)), hi = int(unsigned(b >> 32));
  // rows the move does not write keep `old`: 0 for a sum, the lane's own value for a max
  const int tlo = __builtin_amdgcn_update_dpp(MAX ? lo : 0, lo, CTRL, ROWMASK, 0xF, false);
  const int thi = __builtin_amdgcn_update_dpp(MAX ? hi : 0, hi, CTRL, ROWMASK, 0xF, false);
  const double t =
      __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(unsigned(thi)) << 32) | unsigned(tlo)));
  return MAX ? fmax(v, t) : v + t;
}

template <bool MAX>
__device__ __forceinline__ double waveReduce(double v) {
  v = dppCombine<0xB1, 0xF, MAX>(v);   // quad_perm [1,0,3,2]
  v = dppCombine<0x4E, 0xF, MAX>(v);   // quad_perm [2,3,0,1]
  v = dppCombine<0x124, 0xF, MAX>(v);  // row_ror:4
  v = dppCombine<0x128, 0xF, MAX>(v);  // row_ror:8
  v = dppCombine<0x142, 0xA, MAX>(v);  // row_bcast:15 into rows 1, 3
  v = dppCombine<0x143, 0xC, MAX>(v);  // row_bcast:31 into rows 2, 3
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const int lo = __builtin_amdgcn_readlane(int(unsigned(b)), kWave - 1);
  const int hi = __builtin_amdgcn_readlane(int(unsigned(b >> 32)), kWave - 1);
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(unsigned(hi)) << 32) | unsigned(lo)));
}
__device__ __forceinline__ double waveSum(double v) { return waveReduce<false>(v); }
__device__ __forceinline__ double waveMax(double v) { return waveReduce<true>(v); }

// Fold K values per thread over the block (fixed order); result valid in
// thread 0.  Bit q of MAXMASK selects max instead of sum for value q.
template <int K, unsigned MAXMASK>
__device__ __forceinline__ void blockFold(double (&v)[K], double* sh) {
  const int tid = int(threadIdx.x) + wgSizeX() * (int(threadIdx.y) + wgSizeY() * int(threadIdx.z));
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  const int nw = (wgSizeX() * wgSizeY() * wgSizeZ()) / kWave;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    v[q] = ((MAXMASK >> q) & 1u) ? waveMax(v[q]) : waveSum(v[q]);
    if (lane == 0) sh[q * 16 + wid] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) {
      double a = sh[q * 16];
      for (int w = 1; w < nw; ++w)
        a = ((MAXMASK >> q) & 1u) ? fmax(a, sh[q * 16 + w]) : a + sh[q * 16 + w];
      v[q] = a;
    }
  }
  __syncthreads();
}

// Write this block's K folded values to partials[q * stride + blockLinear].
// Plain stores: the consumer is a later kernel (k_fin-style finalize), so the
// kernel boundary orders them and no fence or atomic is needed.
template <int K, unsigned MAXMASK = 0u>
__device__ __forceinline__ void blockToPartials(double (&v)[K], double* partials, int stride, int blockLinear) {
  __shared__ double sh[K * 16];
  blockFold<K, MAXMASK>(v, sh);
  // the block total sits in linear thread 0 only (every index counts: blocks
  // may be 3-D)
  if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) stA(&partials[q * stride + blockLinear], v[q]);
  }
}

// A thread's strided share of K reductions: a[q] combines ld(q, b) for b = t,
// t + nt, ... < n in that order (bit q of MAXMASK: max instead of sum).  The
// loads of kFoldBatch strides of every row are issued before they are
// combined, so a thread makes one memory round trip per batch instead of one
// per element (a loop that adds each load as it arrives waits for every one).
constexpr int kFoldBatch = 8;
template <int K, unsigned MAXMASK, class LD>
__device__ __forceinline__ void stridedFold(int t, int nt, int n, const LD& ld, double (&a)[K]) {
#pragma unroll
  for (int q = 0; q < K; ++q) a[q] = 0.0;
  for (int b0 = t; b0 < n; b0 += kFoldBatch * nt) {
    double v[K][kFoldBatch];
#pragma unroll
    for (int u = 0; u < kFoldBatch; ++u) {
      const int b = b0 + u * nt;
#pragma unroll
      for (int q = 0; q < K; ++q) v[q][u] = b < n ? ld(q, b) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kFoldBatch; ++u)
      if (b0 + u * nt < n)
#pragma unroll
        for (int q = 0; q < K; ++q) a[q] = ((MAXMASK >> q) & 1u) ? fmax(a[q], v[q][u]) : a[q] + v[q][u];
  }
}

// Fold P partials per value in a fixed order inside one block; the totals are
// returned in every thread.
template <int K, unsigned MAXMASK = 0u>
__device__ __forceinline__ void foldPartials(const double* partials, int stride, int P, double (&out)[K]) {
  __shared__ double sh[K * 16];
  __shared__ double res[K];
  double a[K];
  stridedFold<K, MAXMASK>(int(threadIdx.x), wgSizeX(), P,
                          [&](int q, int b) { return ldA(&partials[q * stride + b]); }, a);
  blockFold<K, MAXMASK>(a, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < K; ++q) res[q] = a[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) out[q] = res[q];
}

constexpr int kFinThreads = 1024;  // threads of a one-block finalize kernel

// XCD-aware tile order.  Workgroups are handed to the 8 XCDs round-robin by
// linear block id, and each XCD has its own L2; renumbering so that XCD x
// receives the contiguous tile range [x*q + min(x, rem), ...) keeps the planes
// a stencil shares with its neighbours inside one L2.  A bijection on [0, B).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcdTile(int b, int B) {
  const int q = B / kXcds, rem = B % kXcds;
  const int x = b % kXcds, idx = b / kXcds;
  return x * q + min(x, rem) + idx;
}
// 3-D block coordinates of this workgroup's tile under xcdTile.
struct Blk3 {
  int x, y, z, lin;
};
__device__ __forceinline__ Blk3 blk3Tiled() {
  const int B = int(gridDim.x * gridDim.y * gridDim.z);
  const int b = int(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
  const int t = xcdTile(b, B);
  Blk3 r;
  r.lin = t;
  r.x = t % int(gridDim.x);
  const int u = t / int(gridDim.x);
  r.y = u % int(gridDim.y);
  r.z = u / int(gridDim.y);
  return r;
}

// Block coordinates of a cell-kernel launch without integer division.
// Workgroups go to the 8 XCDs round-robin by linear id, and each XCD has its
// own L2.  Plain order (slab = 0): the launch coordinates.  XCD slabs
// (slab = 1): the launch is 8 gx x gy x ceil(gz / 8) blocks and block x' runs
// tile x' >> 3 of plane (x' & 7) * gridDim.z + z'; its linear id is x' mod 8,
// so XCD s receives planes [s * gridDim.z, (s + 1) * gridDim.z) in order and
// the k +- 1 planes a stencil shares stay in its L2.  Planes past the grid's
// last fall outside it like any padding.  Row bands (slab = 2): the launch is
// 8 gx x ceil(gy / 8) x gz blocks and block x' runs tile x' >> 3 of block row
// (x' & 7) * gridDim.y + y' in plane z', so XCD s receives one band of block
// rows of every plane, plane by plane: the i and j halos a tile shares with
// its neighbours and the k +- 1 planes stay in one L2 (except on the band's
// two edge rows), while all XCDs stream the same plane at a time.  Rows past
// the grid's last fall outside it.
__device__ __forceinline__ Blk3 blk3(int slab) {
  Blk3 r;
  r.lin = int(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
  if (slab == 1) {
    r.x = int(blockIdx.x >> 3);
    r.y = int(blockIdx.y);
    r.z = int(blockIdx.x & 7u) * int(gridDim.z) + int(blockIdx.z);
  } else if (slab == 2) {
    r.x = int(blockIdx.x >> 3);
    r.y = int(blockIdx.x & 7u) * int(gridDim.y) + int(blockIdx.y);
    r.z = int(blockIdx.z);
  } else {
    r.x = int(blockIdx.x);
    r.y = int(blockIdx.y);
    r.z = int(blockIdx.z);
  }
  return r;
}

}  // namespace nbl
