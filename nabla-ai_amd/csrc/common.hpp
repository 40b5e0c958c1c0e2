// common.hpp — device/host plumbing shared by the gfx950 kernels.
//
// * HIP error handling that turns a failed runtime call into an nbl status and
//   a thread-local message (no exception crosses the C-ABI).
// * Deterministic grid reductions: every block folds its values with a fixed
//   wavefront/LDS tree, writes one partial, and the last block to arrive (ticket
//   counter) folds the partials in index order.  The grid size of a reducing
//   launch is fixed per problem size, so the result is bit-reproducible run to
//   run.  The last block also resets the ticket, so no memset is needed between
//   launches (graph-capture friendly).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace nbl {

constexpr int kBlock = 256;          // threads per block for streaming kernels
constexpr int kWave = 64;            // CDNA wavefront
constexpr int kMaxRedBlocks = 2048;  // cap on blocks of a reducing launch
constexpr int kMaxRedVals = 8;       // values reduced together by one launch

struct HipError : std::runtime_error {
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

// Scratch for deterministic reductions: partial sums per block and a ticket.
struct RedBuf {
  double* partials;   // kMaxRedVals * kMaxRedBlocks
  unsigned* ticket;   // one counter, zero between launches
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double waveSum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ double waveMax(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

// Fold K values per thread over the block (fixed order); result valid in
// thread 0.  Bit q of MAXMASK selects max instead of sum for value q.
template <int K, unsigned MAXMASK>
__device__ __forceinline__ void blockFold(double (&v)[K], double* sh) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    v[q] = ((MAXMASK >> q) & 1u) ? waveMax(v[q]) : waveSum(v[q]);
    if (lane == 0) sh[q * 16 + wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
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

// Grid-wide deterministic reduction of K per-thread values (sum, or max for
// the values flagged in MAXMASK; max operands must be >= 0).  Returns true in
// every thread of the last block to finish; `out` then holds the totals in
// thread 0.  Requires blockDim.x <= 1024 and gridDim.x <= kMaxRedBlocks.
template <int K, unsigned MAXMASK = 0u>
__device__ __forceinline__ bool gridFold(double (&v)[K], RedBuf rb, double (&out)[K]) {
  static_assert(K <= kMaxRedVals, "too many values");
  __shared__ double sh[K * 16];
  __shared__ int last;
  blockFold<K, MAXMASK>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) rb.partials[q * kMaxRedBlocks + blockIdx.x] = v[q];
    __threadfence();
    const unsigned t = atomicAdd(rb.ticket, 1u);
    last = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double a[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    a[q] = 0.0;
    for (int b = threadIdx.x; b < int(gridDim.x); b += blockDim.x) {
      const double pv = __hip_atomic_load(&rb.partials[q * kMaxRedBlocks + b], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
      a[q] = ((MAXMASK >> q) & 1u) ? fmax(a[q], pv) : a[q] + pv;
    }
  }
  blockFold<K, MAXMASK>(a, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) out[q] = a[q];
    *rb.ticket = 0u;
  }
  return true;
}

}  // namespace nbl
