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

// Scratch for deterministic reductions: kMaxRedVals rows of `stride` partials
// (one per block of the largest reducing launch).
struct RedBuf {
  double* partials;
  int stride;
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
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  const int nw = (blockDim.x * blockDim.y * blockDim.z) / kWave;
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
  if (threadIdx.x + threadIdx.y * blockDim.x == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) partials[q * stride + blockLinear] = v[q];
  }
}

// Fold P partials per value in a fixed order inside one block; the totals are
// returned in every thread.
template <int K, unsigned MAXMASK = 0u>
__device__ __forceinline__ void foldPartials(const double* partials, int stride, int P, double (&out)[K]) {
  __shared__ double sh[K * 16];
  __shared__ double res[K];
  double a[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    a[q] = 0.0;
    for (int b = threadIdx.x; b < P; b += blockDim.x) {
      const double pv = partials[q * stride + b];
      a[q] = ((MAXMASK >> q) & 1u) ? fmax(a[q], pv) : a[q] + pv;
    }
  }
  blockFold<K, MAXMASK>(a, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < K; ++q) res[q] = a[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) out[q] = res[q];
}

constexpr int kFinThreads = 1024;  // threads of a one-block finalize kernel

}  // namespace nbl
