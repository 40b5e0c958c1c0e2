// mgkern.hpp — device helpers shared by the multigrid kernel files (mg.hip,
// vcycle.hip): cell coordinates of a level launch, vector loads and stores of
// 1, 2 or 4 consecutive cells, the 7-point operator rebuilt from a cell's
// coupling code in the reference's face order (poisson.cpp:102-134), the
// first damped-Jacobi sweep from zero, and the reductions that a consumer
// kernel folds itself (InFold: alpha, beta, the convergence test; MeanRef:
// pending null-space means).
#pragma once

#include <type_traits>

#include "dev.hpp"

namespace nbl {
namespace {

struct Cell {
  int i, j, k, c, blk;
  bool valid;
};

__device__ __forceinline__ Cell cellOf(int nx, int ny, int nz) {
  Cell q;
  const Blk3 B = blk3(0);
  q.i = B.x * wgSizeX() + int(threadIdx.x);
  q.j = B.y * wgSizeY() + int(threadIdx.y);
  q.k = B.z;
  q.valid = q.i < nx && q.j < ny && q.k < nz;
  q.c = q.i + nx * (q.j + ny * q.k);
  q.blk = B.lin;
  return q;
}

// NC consecutive cells along i per thread (NC = 1, 2 or 4; nx % NC == 0).
// With NC >= 2 every vector access of a cell pair is one 16-byte (FP64) or
// 8-byte (FP32) load or store and the face codes of the group are one 32- or
// 64-bit word.
template <int NC>
struct CellN {
  int i, j, k, c, blk;  // i: first cell of the group; c counts the batch's earlier problems
  int p;                // problem of a batch
  bool valid;
};

// nb > 1: a batch of nb problems whose planes follow each other along the
// launch's z (plane z of the launch is plane z mod nz of problem z / nz).
template <int NC>
__device__ __forceinline__ CellN<NC> cellOfN(int nx, int ny, int nz, int slab, int nb = 1) {
  CellN<NC> q;
  const Blk3 B = blk3(slab);
  q.i = NC * (B.x * wgSizeX() + int(threadIdx.x));
  q.j = B.y * wgSizeY() + int(threadIdx.y);
  q.k = B.z;
  q.p = 0;
  if (nb > 1) {
    q.p = q.k / nz;
    q.k -= q.p * nz;
  }
  q.valid = q.i < nx && q.j < ny && q.k < nz;
  q.c = q.i + nx * (q.j + ny * (q.k + nz * q.p));
  q.blk = B.lin;
  return q;
}

// The first block of this cell kernel's problem (block 0 of a single
// problem): the one that counts the problem's iterations.
template <int NC>
__device__ __forceinline__ bool problemLead(const CellN<NC>& q, int nz, int nb) {
  if (nb <= 1) return blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0;
  return q.blk == q.p * int(gridDim.x * gridDim.y) * nz;
}

// Runs f(std::integral_constant<int, NC>, launch shape) with the level's
// cells per thread.
template <class F>
void withCells(const MgLevel& L, F&& f) {
  if (L.vec == 4) f(std::integral_constant<int, 4>{}, geo3n(L.nx, L.ny, L.nz, 4, L.slab, L.nb));
  else if (L.vec == 2) f(std::integral_constant<int, 2>{}, geo3n(L.nx, L.ny, L.nz, 2, L.slab, L.nb));
  else f(std::integral_constant<int, 1>{}, geo3n(L.nx, L.ny, L.nz, 1, L.slab, L.nb));
}

template <class T>
struct Vec2Of;
template <>
struct Vec2Of<double> {
  using type = double2;
};
template <>
struct Vec2Of<float> {
  using type = float2;
};

template <int NC, class T>
__device__ __forceinline__ void ldN(const T* __restrict__ p, int c, T (&v)[NC]) {
  if constexpr (NC == 1) {
    v[0] = p[c];
  } else {
    using V2 = typename Vec2Of<T>::type;
#pragma unroll
    for (int h = 0; h < NC / 2; ++h) {
      const V2 w = *reinterpret_cast<const V2*>(p + c + 2 * h);
      v[2 * h] = w.x;
      v[2 * h + 1] = w.y;
    }
  }
}
template <int NC, class T>
__device__ __forceinline__ void stN(T* __restrict__ p, int c, const T (&v)[NC]) {
  if constexpr (NC == 1) {
    p[c] = v[0];
  } else {
    using V2 = typename Vec2Of<T>::type;
#pragma unroll
    for (int h = 0; h < NC / 2; ++h) {
      V2 w;
      w.x = v[2 * h];
      w.y = v[2 * h + 1];
      *reinterpret_cast<V2*>(p + c + 2 * h) = w;
    }
  }
}
template <int NC>
__device__ __forceinline__ void ldCodes(const uint16_t* __restrict__ code, int c, uint16_t (&v)[NC]) {
  if constexpr (NC == 1) {
    v[0] = code[c];
  } else if constexpr (NC == 2) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(code + c);
    v[0] = uint16_t(w & 0xffffu);
    v[1] = uint16_t(w >> 16);
  } else {
    const uint2 w = *reinterpret_cast<const uint2*>(code + c);
    v[0] = uint16_t(w.x & 0xffffu);
    v[1] = uint16_t(w.x >> 16);
    v[2] = uint16_t(w.y & 0xffffu);
    v[3] = uint16_t(w.y >> 16);
  }
}

// Neighbour across face `dir` (periodic wrap where the code says kind 1 at a
// domain face): flat index and coordinates.
__device__ __forceinline__ int nbOf(const MgLevel& L, int c, int i, int j, int k, int dir, int& ii, int& jj, int& kk) {
  ii = i;
  jj = j;
  kk = k;
  const int sy = L.nx, sz = L.nx * L.ny;
  switch (dir) {
    case 0: if (i > 0) { ii = i - 1; return c - 1; } ii = L.nx - 1; return c + L.nx - 1;
    case 1: if (i < L.nx - 1) { ii = i + 1; return c + 1; } ii = 0; return c - (L.nx - 1);
    case 2: if (j > 0) { jj = j - 1; return c - sy; } jj = L.ny - 1; return c + (L.ny - 1) * sy;
    case 3: if (j < L.ny - 1) { jj = j + 1; return c + sy; } jj = 0; return c - (L.ny - 1) * sy;
    case 4: if (k > 0) { kk = k - 1; return c - sz; } kk = L.nz - 1; return c + (L.nz - 1) * sz;
    default: if (k < L.nz - 1) { kk = k + 1; return c + sz; } kk = 0; return c - (L.nz - 1) * sz;
  }
}

// Value of x at a fluid cell as the reference holds it at that point of the
// cycle: x - pending mean (+ the parent's coarse correction after prolongation,
// PROLONG; the parent of (i,j,k) is (i>>sx, j>>sy, k>>sz), factors being 1 or 2).
template <class T, bool PROLONG = false>
struct XView {
  const T* __restrict__ x;
  T m;
  bool hasM;
  const T* __restrict__ ec;  // PROLONG: add the parent's correction ec[parent] - em
  int ncx, ncy, sx, sy, sz;
  T em;
  bool hasEm;
  __device__ __forceinline__ T fix(T v, int i, int j, int k) const {
    if (hasM) v = v - m;
    if constexpr (PROLONG) {
      T e = ec[(i >> sx) + ncx * ((j >> sy) + ncy * (k >> sz))];
      if (hasEm) e = e - em;
      v = v + e;
    }
    return v;
  }
  __device__ __forceinline__ T at(int c, int i, int j, int k) const { return fix(x[c], i, j, k); }
  template <int NC>
  __device__ __forceinline__ void pack(int c, int i, int j, int k, T (&v)[NC]) const {
    ldN<NC>(x, c, v);
#pragma unroll
    for (int e = 0; e < NC; ++e) v[e] = fix(v[e], i + e, j, k);
  }
};

template <class T>
__device__ __forceinline__ XView<T> plainView(const T* x) {
  return XView<T>{x, T(0), false, nullptr, 1, 1, 0, 0, 0, T(0), false};
}

// (A x)_c from the centre value and the six neighbour values (order -x, +x,
// -y, +y, -z, +z) and the diagonal sum dg: the couplings are accumulated in
// the reference's face order (poisson.cpp:102-134).
template <class T, bool ZERO>
__device__ __forceinline__ T applyDecode(const MgLevel& L, uint16_t code, T xc, const T (&xn)[6], T& dg) {
  const T h2[3] = {T(L.h2[0]), T(L.h2[1]), T(L.h2[2])};
  T d = T(0), off = T(0);
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    const int kd = kindOf(code, dir);
    const int a = dir >> 1;
    if (kd == 1) {
      d += h2[a];
      off += h2[a] * (ZERO ? T(0) : xn[dir]);
    } else if (kd == 2) {
      d += T(2) * h2[a];
    }
  }
  dg = d;
  return d * xc - off;
}

// (A x)_c for a fluid cell and the diagonal sum dg (poisson.cpp:102-134): the
// couplings are accumulated in the reference's face order.  The six neighbour
// values are read before the face kinds are decoded (every neighbour index is
// in range: nbOf wraps), so a wavefront issues all its stencil loads in one
// round trip instead of waiting for the face code first.
template <class T, bool ZERO, class XV>
__device__ __forceinline__ T applyRow(const MgLevel& L, uint16_t code, int c, int i, int j, int k, const XV& X,
                                      T& dg) {
  T xn[6];
  T xc = T(0);
  if constexpr (!ZERO) {
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      int ii, jj, kk;
      const int nb = nbOf(L, c, i, j, k, dir, ii, jj, kk);
      xn[dir] = X.at(nb, ii, jj, kk);
    }
    xc = X.at(c, i, j, k);
  }
  return applyDecode<T, ZERO>(L, code, xc, xn, dg);
}

// ---------------------------------------------------------------------------
// Smoother: x <- x + 0.8 (b - A x)/diag on fluid cells, 0 elsewhere
// (multigrid.cpp:102-114), with the fused prolongation of the parent
// correction (multigrid.cpp:158-173) in SW_PROLONG, and per-block fluid sums
// of the result for the projection that follows a sweep on singular systems.
// Pending means (`mb`, `mx`, `mec`) are device slots, read here.
// ---------------------------------------------------------------------------
enum SweepMode : int { SW_PLAIN = 0, SW_ZERO = 1, SW_PROLONG = 2 };

// First sweep from zero at one cell, in the expression order of k_sweep.
template <class T>
__device__ __forceinline__ T firstSweep(const MgLevel& L, uint16_t code, int c, int i, int j, int k, T bv) {
  if (!(code & kFluidBit)) return T(0);
  T dg;
  const T Ax = applyRow<T, true>(L, code, c, i, j, k, plainView<T>(nullptr), dg);
  const T diag = dg > T(0) ? dg : T(1);
  return T(0) + T(0.8) * (bv - Ax) / diag;
}

// Centre values and the six neighbour values (order -x, +x, -y, +y, -z, +z)
// of the NC cells of a group, read through a view X (plain, pending mean,
// prolongation, ...).  Neighbour indices wrap periodically like nbOf; the
// group's inner neighbours along i come from the centre pack.
template <class T, int NC, class XV>
__device__ __forceinline__ void gatherN(const MgLevel& L, const XV& X, int c, int i, int j, int k, T (&xc)[NC],
                                        T (&xn)[NC][6]) {
  const int sy = L.nx, sz = L.nx * L.ny;
  T jm[NC], jp[NC], km[NC], kp[NC];
  X.template pack<NC>(c, i, j, k, xc);
  const int im = i > 0 ? i - 1 : L.nx - 1;
  const int cm = i > 0 ? c - 1 : c + L.nx - 1;
  const int ip = i + NC < L.nx ? i + NC : 0;
  const int cp = i + NC < L.nx ? c + NC : c + NC - L.nx;
  const T xm = X.at(cm, im, j, k);
  const T xp = X.at(cp, ip, j, k);
  const int jjm = j > 0 ? j - 1 : L.ny - 1, jjp = j < L.ny - 1 ? j + 1 : 0;
  const int kkm = k > 0 ? k - 1 : L.nz - 1, kkp = k < L.nz - 1 ? k + 1 : 0;
  X.template pack<NC>(j > 0 ? c - sy : c + (L.ny - 1) * sy, i, jjm, k, jm);
  X.template pack<NC>(j < L.ny - 1 ? c + sy : c - (L.ny - 1) * sy, i, jjp, k, jp);
  X.template pack<NC>(k > 0 ? c - sz : c + (L.nz - 1) * sz, i, j, kkm, km);
  X.template pack<NC>(k < L.nz - 1 ? c + sz : c - (L.nz - 1) * sz, i, j, kkp, kp);
#pragma unroll
  for (int e = 0; e < NC; ++e) {
    xn[e][0] = e == 0 ? xm : xc[e - 1];
    xn[e][1] = e == NC - 1 ? xp : xc[e + 1];
    xn[e][2] = jm[e];
    xn[e][3] = jp[e];
    xn[e][4] = km[e];
    xn[e][5] = kp[e];
  }
}

// ---------------------------------------------------------------------------
// Finalizes folded into the consuming kernel (InFold).  Every block reads the
// previous kernel's partials with plain loads (the kernel boundary orders
// them) and folds them in one fixed order.  A kernel whose own block 0 may
// stop the solve reads `done` once per block, so that every thread of a block
// takes the same branch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int linearTid() {
  return int(threadIdx.x) + wgSizeX() * (int(threadIdx.y) + wgSizeY() * int(threadIdx.z));
}
__device__ __forceinline__ bool firstBlock() { return blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0; }

template <int K>
__device__ __forceinline__ void foldInline(const InFold& f, double (&out)[K]) {
  __shared__ double sh[K * 16];
  __shared__ double res[K];
  const int tid = linearTid(), nt = wgSizeX() * wgSizeY() * wgSizeZ();
  double a[K];
  stridedFold<K, 0u>(tid, nt, f.n, [&](int q, int b) { return f.parts[q * f.stride + b]; }, a);
  blockFold<K, 0u>(a, sh);
  if (tid == 0)
    for (int q = 0; q < K; ++q) res[q] = a[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) out[q] = res[q];
}

// The solve's stop flag (ctl may be null: never stopped) and the values of K
// pending means, in every thread (block-wide: every thread of the block calls
// it).  Folded means share one pass over their partials, in the same memory
// round trip as the flag (thread 0 reads it).  True: the solve has stopped.
template <int K>
__device__ __forceinline__ bool stopOrMeans(const SolveCtl* ctl, const MeanRef (&m)[K], double (&v)[K]) {
  bool fold = false;
#pragma unroll
  for (int q = 0; q < K; ++q) fold = fold || m[q].parts != nullptr;
  if (!fold) {
    if (ctl && ldA(&ctl->done)) return true;
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = m[q].slot ? ldA(m[q].slot) : 0.0;
    return false;
  }
  __shared__ double sh[K * 16];
  __shared__ double res[K];
  __shared__ int st;
  const int tid = linearTid(), nt = wgSizeX() * wgSizeY() * wgSizeZ();
  const int d = (tid == 0 && ctl) ? ldA(&ctl->done) : 0;
  double a[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s[1] = {0.0};
    if (m[q].parts) {
      const double* pp = m[q].parts;
      stridedFold<1, 0u>(tid, nt, m[q].n, [&](int, int b) { return pp[b]; }, s);
    }
    a[q] = s[0];
  }
  blockFold<K, 0u>(a, sh);
  if (tid == 0) {
    for (int q = 0; q < K; ++q) res[q] = a[q];
    st = d;
  }
  __syncthreads();
  if (st) return true;
#pragma unroll
  for (int q = 0; q < K; ++q) a[q] = res[q];
#pragma unroll
  for (int q = 0; q < K; ++q)
    v[q] = m[q].parts ? (m[q].nf > 0.0 ? a[q] / m[q].nf : 0.0) : (m[q].slot ? ldA(m[q].slot) : 0.0);
  return false;
}

// The partials of an InFold folded together with the solve's stop flag and
// up to two SolveCtl scalars (e0, e1; may be null) in one memory round trip:
// thread 0 reads the scalars while every thread reads its share of the
// partials.  Every thread gets the totals, the scalars and a block-uniform
// stop flag.
template <int K>
__device__ __forceinline__ void foldStop(const InFold& f, const SolveCtl* ctl, const double* e0, const double* e1,
                                         double (&out)[K], double (&ex)[2], bool& stop) {
  __shared__ double sh[K * 16];
  __shared__ double res[K + 2];
  __shared__ int st;
  const int tid = linearTid(), nt = wgSizeX() * wgSizeY() * wgSizeZ();
  int d = 0;
  double x0 = 0.0, x1 = 0.0;
  if (tid == 0) {
    d = ldA(&ctl->done);
    if (e0) x0 = ldA(e0);
    if (e1) x1 = ldA(e1);
  }
  double a[K];
  stridedFold<K, 0u>(tid, nt, f.n, [&](int q, int b) { return f.parts[q * f.stride + b]; }, a);
  blockFold<K, 0u>(a, sh);
  if (tid == 0) {
    for (int q = 0; q < K; ++q) res[q] = a[q];
    res[K] = x0;
    res[K + 1] = x1;
    st = d;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) out[q] = res[q];
  ex[0] = res[K];
  ex[1] = res[K + 1];
  stop = st != 0;
}

// FIN_CONV: |r_new| <= tol |b| stops the solve.  True when this kernel must
// do nothing (stopped before or now).
__device__ __forceinline__ bool inlineConv(const InFold& f, SolveCtl* ctl) {
  double t[1], ex[2];
  bool stop;
  foldStop<1>(f, ctl, &ctl->tol, &ctl->bnorm, t, ex, stop);
  if (stop) return true;
  const bool conv = sqrt(t[0]) <= ex[0] * ex[1];
  if (firstBlock() && linearTid() == 0) {
    stA(&ctl->rr, t[0]);
    if (conv) {
      stA(&ctl->it, ldA(&ctl->it) + 1);
      stA(&ctl->converged, 1);
      stA(&ctl->done, 1);
    }
  }
  return conv;
}

// FIN_ALPHA: alpha = rz / p.Ap; a vanishing p.Ap stops the solve (breakdown).
// False when this kernel must do nothing.
__device__ __forceinline__ bool inlineAlpha(const InFold& f, SolveCtl* ctl, double& alpha) {
  double t[1], ex[2];
  bool stop;
  foldStop<1>(f, ctl, &ctl->rzq[f.k & 1], nullptr, t, ex, stop);
  if (stop) return false;
  const bool w = firstBlock() && linearTid() == 0;
  if (fabs(t[0]) < 1e-300) {
    if (w) {
      stA(&ctl->pAp, t[0]);
      stA(&ctl->broke, 1);
      stA(&ctl->done, 1);
    }
    return false;
  }
  alpha = ex[0] / t[0];
  if (w) {
    stA(&ctl->pAp, t[0]);
    stA(&ctl->alpha, alpha);
  }
  return true;
}

// FIN_BETA_FLEX: beta = (z.r_new - z.r) / rz_prev, rz = z.r_new.  False when
// the solve has stopped.
__device__ __forceinline__ bool inlineBeta(const InFold& f, SolveCtl* ctl, double& be) {
  double t[2], ex[2];
  bool stop;
  foldStop<2>(f, ctl, &ctl->rzq[(f.k - 1) & 1], nullptr, t, ex, stop);
  if (stop) return false;
  be = (t[0] - t[1]) / ex[0];
  if (firstBlock() && linearTid() == 0) {
    stA(&ctl->beta, be);
    stA(&ctl->rz, t[0]);
    stA(&ctl->rzq[f.k & 1], t[0]);
  }
  return true;
}

// FIN_BETA_FLEX with z's fluid sum folded beside the dots in a third row (z's
// mean pending): m = sum / nf, the mean the projection would remove.
__device__ __forceinline__ bool inlineBetaMean(const InFold& f, SolveCtl* ctl, double nf, double& be, double& m) {
  double t[3], ex[2];
  bool stop;
  foldStop<3>(f, ctl, &ctl->rzq[(f.k - 1) & 1], nullptr, t, ex, stop);
  if (stop) return false;
  be = (t[0] - t[1]) / ex[0];
  m = t[2] / nf;
  if (firstBlock() && linearTid() == 0) {
    stA(&ctl->beta, be);
    stA(&ctl->rz, t[0]);
    stA(&ctl->rzq[f.k & 1], t[0]);
  }
  return true;
}

}  // namespace
}  // namespace nbl
