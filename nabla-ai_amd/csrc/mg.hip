// mg.hip — gfx950 kernels of the pressure Poisson solve
// (reference core/src/flow/poisson.cpp and core/src/flow/multigrid.cpp).
//
// Memory model: one compact array per vector and level (i fastest) plus a
// 16-bit per-cell coupling code; the 7-point operator, its diagonal and the
// periodic wrap are all rebuilt from the code, so a sweep moves 3 vectors and
// 2 bytes per cell.  Cell kernels run on a 3-D launch (64..8 threads along i,
// 4..32 rows of j per block, one k per grid z), so (i, j, k) come from the
// launch coordinates with no integer division and loads along i coalesce.
//
// The reference's null-space projection ("removeMean", multigrid.cpp:81-99)
// is applied lazily: the kernel that produces a vector also writes per-block
// partial sums of it, a one-block finalize kernel turns them into the mean,
// and every consumer subtracts the mean when it reads the vector.  x - mean is
// the same rounded value whether it is stored first or formed on the fly, so
// this is bit-identical per element to the reference.
//
// Every kernel of a flexible-CG iteration first checks ctl->done, so the host
// can enqueue iterations before it knows whether the solve has converged.
#include "dev.hpp"

namespace nbl {

Geo3 geo3(int nx, int ny, int nz) {
  const int bx = nx >= 64 ? 64 : nx >= 32 ? 32 : nx >= 16 ? 16 : 8;
  const int by = kBlock / bx;
  Geo3 g;
  g.block = dim3(bx, by, 1);
  g.grid = dim3((nx + bx - 1) / bx, (ny + by - 1) / by, nz);
  g.blocks = int(g.grid.x * g.grid.y * g.grid.z);
  return g;
}

namespace {

struct Cell {
  int i, j, k, c, blk;
  bool valid;
};

__device__ __forceinline__ Cell cellOf(int nx, int ny, int nz) {
  Cell q;
  const Blk3 B = blk3();
  q.i = B.x * blockDim.x + threadIdx.x;
  q.j = B.y * blockDim.y + threadIdx.y;
  q.k = B.z;
  q.valid = q.i < nx && q.j < ny && q.k < nz;
  q.c = q.i + nx * (q.j + ny * q.k);
  q.blk = B.lin;
  return q;
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
  __device__ __forceinline__ T at(int c, int i, int j, int k) const {
    T v = x[c];
    if (hasM) v = v - m;
    if constexpr (PROLONG) {
      T e = ec[(i >> sx) + ncx * ((j >> sy) + ncy * (k >> sz))];
      if (hasEm) e = e - em;
      v = v + e;
    }
    return v;
  }
};

template <class T>
__device__ __forceinline__ XView<T> plainView(const T* x) {
  return XView<T>{x, T(0), false, nullptr, 1, 1, 0, 0, 0, T(0), false};
}

// (A x)_c for a fluid cell and the diagonal sum dg (poisson.cpp:102-134): the
// couplings are accumulated in the reference's face order.  The six neighbour
// values are read before the face kinds are decoded (every neighbour index is
// in range: nbOf wraps), so a wavefront issues all its stencil loads in one
// round trip instead of waiting for the face code first.
template <class T, bool ZERO, class XV>
__device__ __forceinline__ T applyRow(const MgLevel& L, uint16_t code, int c, int i, int j, int k, const XV& X,
                                      T& dg) {
  const T h2[3] = {T(L.h2[0]), T(L.h2[1]), T(L.h2[2])};
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

// ---------------------------------------------------------------------------
// Smoother: x <- x + 0.8 (b - A x)/diag on fluid cells, 0 elsewhere
// (multigrid.cpp:102-114), with the fused prolongation of the parent
// correction (multigrid.cpp:158-173) in SW_PROLONG, and per-block fluid sums
// of the result for the projection that follows a sweep on singular systems.
// Pending means (`mb`, `mx`, `mec`) are device slots, read here.
// SW_PAIR runs the first two sweeps from a zero guess in one pass (systems
// without the projection between sweeps): the first sweep is pointwise,
// x1 = 0 + 0.8 (b - 0)/diag, so each cell re-evaluates it at its stencil points.
// ---------------------------------------------------------------------------
enum SweepMode : int { SW_PLAIN = 0, SW_ZERO = 1, SW_PROLONG = 2, SW_PAIR = 3 };

// First sweep from zero at one cell, in the expression order of k_sweep.
template <class T>
__device__ __forceinline__ T firstSweep(const MgLevel& L, uint16_t code, int c, int i, int j, int k, T bv) {
  if (!(code & kFluidBit)) return T(0);
  T dg;
  const T Ax = applyRow<T, true>(L, code, c, i, j, k, plainView<T>(nullptr), dg);
  const T diag = dg > T(0) ? dg : T(1);
  return T(0) + T(0.8) * (bv - Ax) / diag;
}

template <class T>
struct FirstSweepView {
  const MgLevel* L;
  const T* __restrict__ b;
  __device__ __forceinline__ T at(int c, int i, int j, int k) const {
    return firstSweep<T>(*L, L->code[c], c, i, j, k, b[c]);
  }
};

template <class T, int MODE>
__global__ void __launch_bounds__(kBlock) k_sweep(MgLevel L, const T* __restrict__ b, const double* mb,
                                                  const T* __restrict__ x, const double* mx, const T* ec,
                                                  const double* mec, int ncx, int ncy, int sx, int sy, int sz,
                                                  T* __restrict__ xo, int fold, RedBuf rb, const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    const uint16_t code = L.code[q.c];
    if (!(code & kFluidBit)) {
      xo[q.c] = T(0);
    } else {
      T dg, Ax, xc;
      if constexpr (MODE == SW_PAIR) {
        const FirstSweepView<T> X{&L, b};
        Ax = applyRow<T, false>(L, code, q.c, q.i, q.j, q.k, X, dg);
        xc = X.at(q.c, q.i, q.j, q.k);
      } else if constexpr (MODE == SW_PROLONG) {
        const XView<T, true> X{x, mx ? T(ldA(mx)) : T(0), mx != nullptr, ec, ncx, ncy, sx, sy, sz,
                               mec ? T(ldA(mec)) : T(0), mec != nullptr};
        Ax = applyRow<T, false>(L, code, q.c, q.i, q.j, q.k, X, dg);
        xc = X.at(q.c, q.i, q.j, q.k);
      } else {
        const XView<T> X{x, mx ? T(ldA(mx)) : T(0), mx != nullptr, nullptr, 1, 1, 0, 0, 0, T(0), false};
        Ax = applyRow<T, MODE == SW_ZERO>(L, code, q.c, q.i, q.j, q.k, X, dg);
        xc = MODE == SW_ZERO ? T(0) : X.at(q.c, q.i, q.j, q.k);
      }
      const T diag = dg > T(0) ? dg : T(1);
      T bv = b[q.c];
      if (mb) bv = bv - T(ldA(mb));
      const T xn = xc + T(0.8) * (bv - Ax) / diag;
      xo[q.c] = xn;
      acc[0] = double(xn);
    }
  }
  if (fold) blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// Residual r = b - A x on fine fluid cells restricted to the coarse grid as
// the mean over fluid children (multigrid.cpp:116-154).  One thread per fine
// cell; a block covers whole coarse cells (blockDim.z = fz, even x/y extents),
// stages its residuals in LDS, and the thread at each coarse cell's first
// child sums the children in the reference's k, j, i order.  ZEROX: x == 0
// (restriction of b itself).
template <class T, bool ZEROX>
__global__ void __launch_bounds__(kBlock) k_resid_restrict(MgLevel L, MgLevel Lc, const T* __restrict__ b,
                                                           const double* mb, const T* __restrict__ x,
                                                           const double* mx, T* __restrict__ rc, int fold,
                                                           RedBuf rb, const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  __shared__ T rs[kBlock];
  __shared__ unsigned char fls[kBlock];
  const XView<T> X{x, mx ? T(ldA(mx)) : T(0), mx != nullptr, nullptr, 1, 1, 0, 0, 0, T(0), false};
  const Blk3 B = blk3();
  const int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
  const int bx = blockDim.x, by = blockDim.y;
  const int i = B.x * bx + tx, j = B.y * by + ty, k = B.z * int(blockDim.z) + tz;
  const int tid = tx + bx * (ty + by * tz);
  T r = T(0);
  bool fl = false;
  if (i < L.nx && j < L.ny && k < L.nz) {
    const int c = i + L.nx * (j + L.ny * k);
    const uint16_t code = L.code[c];
    if (code & kFluidBit) {
      fl = true;
      T dg;
      const T Ax = applyRow<T, ZEROX>(L, code, c, i, j, k, X, dg);
      T bv = b[c];
      if (mb) bv = bv - T(ldA(mb));
      r = bv - Ax;
    }
  }
  rs[tid] = r;
  fls[tid] = fl ? 1 : 0;
  __syncthreads();
  double acc[1] = {0.0};
  if (tz == 0 && tx % Lc.fx == 0 && ty % Lc.fy == 0) {
    const int I = i / Lc.fx, J = j / Lc.fy, K = k / Lc.fz;
    if (I < Lc.nx && J < Lc.ny && K < Lc.nz) {
      const int C = I + Lc.nx * (J + Lc.ny * K);
      if (!(Lc.code[C] & kFluidBit)) {
        rc[C] = T(0);
      } else {
        T s = T(0);
        int cnt = 0;
        for (int dk = 0; dk < Lc.fz; ++dk)
          for (int dj = 0; dj < Lc.fy; ++dj)
            for (int di = 0; di < Lc.fx; ++di) {
              const int t = (tx + di) + bx * ((ty + dj) + by * dk);
              if (fls[t]) {
                s += rs[t];
                ++cnt;
              }
            }
        const T rv = s / T(cnt);
        rc[C] = rv;
        acc[0] = double(rv);
      }
    }
  }
  if (fold) blockToPartials<1>(acc, rb.partials, rb.stride, B.lin);
}

// ---------------------------------------------------------------------------
// Coarsest level: Jacobi-PCG to 1e-6 in at most 500 iterations, as
// CgPoisson(1e-6, 500).solve (multigrid.cpp:178-183, poisson.cpp:136-240).
// One workgroup of 256 threads (one wavefront per SIMD: the per-wavefront
// scalar work of an iteration is then done once per SIMD), up to 8 cells per
// thread, every vector in
// registers and the preconditioned residual u mirrored in LDS for the
// stencil (absent neighbours read a zero slot).  The iteration is the
// Chronopoulos-Gear form of PCG (s = A p carried by recurrence), which needs a
// single fused reduction of (r.u, Au.u, r.r) and two barriers per iteration;
// in exact arithmetic its iterates are the reference's.  The coarse solve's
// dot products are tree-ordered, so it matches the reference's CG to rounding,
// not bit for bit, and u = r * (1/diag) is used for the same reason.
// ---------------------------------------------------------------------------
constexpr int kCoarseT = 256;
constexpr int kCPT = 8;
constexpr int kCoarseMax = kCoarseT * kCPT;  // the hierarchy stops at <= 2048 cells

template <int K>
__device__ __forceinline__ void coarseSum(double (&v)[K], double (*red)[kCoarseT / kWave]) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    v[q] = waveSum(v[q]);
    if (lane == 0) red[q][wid] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kCoarseT / kWave; ++w) s += red[q][w];
    v[q] = s;
  }
}

__device__ __forceinline__ double fmaT(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return fmaf(a, b, c); }

template <class T>
__global__ void __launch_bounds__(kCoarseT) k_coarse_lds(MgLevel L, int singular, const T* __restrict__ b,
                                                         const double* mb, T* __restrict__ xout,
                                                         const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  __shared__ T ul[kCoarseMax + 1];  // [kCoarseMax] stays 0
  __shared__ double red1[1][kCoarseT / kWave];
  __shared__ double red3[3][kCoarseT / kWave];
  const int tid = threadIdx.x;
  const int N = int(L.n);
  const T h2[3] = {T(L.h2[0]), T(L.h2[1]), T(L.h2[2])};
  T x[kCPT], r[kCPT], p[kCPT], sv[kCPT], inv[kCPT], dsum[kCPT];
  int nb[kCPT][6];
  bool fl[kCPT], live[kCPT];
  const T bm = mb ? T(ldA(mb)) : T(0);
  if (tid == 0) ul[kCoarseMax] = T(0);
  double s0[1] = {0.0};
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const int c = tid + q * kCoarseT;
    live[q] = c < N;
    fl[q] = false;
    x[q] = r[q] = p[q] = sv[q] = dsum[q] = T(0);
    inv[q] = T(1);
#pragma unroll
    for (int d = 0; d < 6; ++d) nb[q][d] = kCoarseMax;
    if (!live[q]) continue;
    const uint16_t code = L.code[c];
    fl[q] = code & kFluidBit;
    T v = b[c];
    if (mb && fl[q]) v = v - bm;
    r[q] = v;
    if (!fl[q]) continue;
    s0[0] += double(v);
    const int i = c % L.nx, j = (c / L.nx) % L.ny, k = c / (L.nx * L.ny);
    T dg = T(0);
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      const int kd = kindOf(code, d);
      if (kd == 1) {
        int ii, jj, kk;
        nb[q][d] = nbOf(L, c, i, j, k, d, ii, jj, kk);
        dg += h2[d >> 1];
      } else if (kd == 2) {
        dg += T(2) * h2[d >> 1];
      }
    }
    dsum[q] = dg;
    inv[q] = T(1) / (dg > T(0) ? dg : T(1));
  }
  if (singular && L.nf > 0) {
    coarseSum<1>(s0, red1);
    const T m = T(s0[0] / double(L.nf));
#pragma unroll
    for (int q = 0; q < kCPT; ++q) r[q] = fl[q] ? r[q] - m : T(0);
  }
  double bn = 0.0, bn2 = 0.0, gammaOld = 1.0, alphaOld = 1.0;
  for (int it = 0;; ++it) {
    // u = M^-1 r to LDS; w = A u; (r.u, w.u, r.r)
    T u[kCPT];
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
      u[q] = r[q] * inv[q];
      if (live[q]) ul[tid + q * kCoarseT] = u[q];
    }
    __syncthreads();
    T w[kCPT];
    double d3[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
      w[q] = T(0);
      if (!live[q]) continue;
      if (fl[q]) {
        T off = h2[0] * (ul[nb[q][0]] + ul[nb[q][1]]);
        off = fmaT(h2[1], ul[nb[q][2]] + ul[nb[q][3]], off);
        off = fmaT(h2[2], ul[nb[q][4]] + ul[nb[q][5]], off);
        w[q] = fmaT(dsum[q], u[q], -off);
      } else {
        w[q] = u[q];
      }
      d3[0] = fma(double(r[q]), double(u[q]), d3[0]);
      d3[1] = fma(double(w[q]), double(u[q]), d3[1]);
      d3[2] = fma(double(r[q]), double(r[q]), d3[2]);
    }
    coarseSum<3>(d3, red3);
    if (it == 0) {
      bn = sqrt(d3[2]) + 1e-300;
      bn2 = (1e-6 * bn) * (1e-6 * bn);
    } else if (d3[2] <= bn2) {
      break;  // |r_it| <= 1e-6 |b| after update it
    }
    if (it == 500) break;
    const double gamma = d3[0];
    const double beta = it == 0 ? 0.0 : gamma / gammaOld;
    const double den = it == 0 ? d3[1] : d3[1] - beta * gamma / alphaOld;  // = p.Ap
    if (fabs(den) < 1e-300) break;
    const double alpha = gamma / den;
    const T a = T(alpha), be = T(beta);
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
      p[q] = fmaT(be, p[q], u[q]);
      sv[q] = fmaT(be, sv[q], w[q]);
      x[q] = fmaT(a, p[q], x[q]);
      r[q] = fmaT(-a, sv[q], r[q]);
    }
    gammaOld = gamma;
    alphaOld = alpha;
  }
  __syncthreads();  // a break right after the reduction leaves red3 being read
  if (singular && L.nf > 0) {
    double sx[1] = {0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q)
      if (fl[q]) sx[0] += double(x[q]);
    coarseSum<1>(sx, red1);
    const T m = T(sx[0] / double(L.nf));
#pragma unroll
    for (int q = 0; q < kCPT; ++q)
      if (fl[q]) x[q] = x[q] - m;
  }
#pragma unroll
  for (int q = 0; q < kCPT; ++q)
    if (live[q]) xout[tid + q * kCoarseT] = x[q];
}

// ---------------------------------------------------------------------------
// Direct coarsest-level solve.  The level's operator (identity rows for solid
// cells; on singular systems the fluid block plus alpha*11^T) is inverted once
// by Gauss-Jordan elimination on the device, projected with the fluid
// mean-removal P on both sides when singular (P A^+ P: the minimum-norm,
// mean-free solution that CG from zero followed by removeMean reaches), and
// each coarse solve is then one dense matrix-vector product.
// ---------------------------------------------------------------------------
// Step k of Gauss-Jordan on the n x 2n augmented matrix G (row-major): the
// pivot row divided by the pivot, and column k, are copied out first so the
// update kernel has no read/write overlap.
__global__ void k_gj_copy(const double* __restrict__ G, int n, int k, double* __restrict__ colk,
                          double* __restrict__ rowk, int* fail, double tiny) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = 2 * n;
  const double piv = G[size_t(k) * w + k];
  if (t == 0 && !(piv > tiny)) *fail = 1;
  if (t < w) rowk[t] = G[size_t(k) * w + t] / piv;
  if (t < n) colk[t] = G[size_t(t) * w + k];
}

__global__ void k_gj_update(double* __restrict__ G, int n, int k, const double* __restrict__ colk,
                            const double* __restrict__ rowk) {
  const int j = k + blockIdx.x * blockDim.x + threadIdx.x;  // columns left of k are settled
  const int i = blockIdx.y;
  const int w = 2 * n;
  if (j >= w) return;
  double* g = G + size_t(i) * w + j;
  *g = i == k ? rowk[j] : *g - colk[i] * rowk[j];
}

// Row means of the inverse over fluid columns (one block per row), for P M P.
__global__ void __launch_bounds__(kBlock) k_inv_rowmean(const double* __restrict__ G, int n,
                                                        const uint16_t* __restrict__ code, long long nf,
                                                        double* __restrict__ rm) {
  const int i = blockIdx.x;
  const double* row = G + size_t(i) * 2 * n + n;
  double a[1] = {0.0};
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    if (code[j] & kFluidBit) a[0] += row[j];
  __shared__ double sh[16];
  blockFold<1, 0u>(a, sh);
  if (threadIdx.x == 0) rm[i] = (code[i] & kFluidBit) ? a[0] / double(nf) : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_inv_total(const double* __restrict__ rm, int n,
                                                      const uint16_t* __restrict__ code, long long nf,
                                                      double* __restrict__ tot) {
  double a[1] = {0.0};
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    if (code[j] & kFluidBit) a[0] += rm[j];
  __shared__ double sh[16];
  blockFold<1, 0u>(a, sh);
  if (threadIdx.x == 0) *tot = a[0] / double(nf);
}

// Final coarse inverse (n x n, FP64 and FP32 copies) from the right half of G.
__global__ void k_inv_store(const double* __restrict__ G, int n, const uint16_t* __restrict__ code, int singular,
                            const double* __restrict__ rm, const double* __restrict__ tot,
                            double* __restrict__ M, float* __restrict__ Mf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  double v = G[size_t(i) * 2 * n + n + j];
  if (singular) {
    const bool fi = code[i] & kFluidBit, fj = code[j] & kFluidBit;
    v = (fi && fj) ? ((v - rm[i]) - rm[j]) + *tot : 0.0;
  }
  M[size_t(i) * n + j] = v;
  Mf[size_t(i) * n + j] = float(v);
}

// x = Minv (b - mb on fluid cells): kDirectRows rows per block, each thread
// striding the columns; fixed-order block fold.
constexpr int kDirectRows = 4;
template <class T, class MT>
__global__ void __launch_bounds__(kBlock) k_coarse_direct(MgLevel L, const MT* __restrict__ M,
                                                          const T* __restrict__ b, const double* mb,
                                                          T* __restrict__ x, const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  const int n = int(L.n);
  const int r0 = blockIdx.x * kDirectRows;
  const double m = mb ? ldA(mb) : 0.0;
  double acc[kDirectRows];
#pragma unroll
  for (int r = 0; r < kDirectRows; ++r) acc[r] = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    T bj = b[j];
    if (mb && (L.code[j] & kFluidBit)) bj = bj - T(m);
    const double bv = double(bj);
#pragma unroll
    for (int r = 0; r < kDirectRows; ++r)
      if (r0 + r < n) acc[r] = fma(double(M[size_t(r0 + r) * n + j]), bv, acc[r]);
  }
  __shared__ double sh[kDirectRows * 16];
  blockFold<kDirectRows, 0u>(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < kDirectRows; ++r)
      if (r0 + r < n) x[r0 + r] = T(acc[r]);
}

// Same solve for coarsest levels above kCoarseMax cells (grids that stop
// coarsening early): vectors in global memory, one workgroup.
template <int K>
__device__ __forceinline__ void blockAll(double (&v)[K], double* sh, double* bc) {
  blockFold<K, 0u>(v, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < K; ++q) bc[q] = v[q];
  __syncthreads();
  for (int q = 0; q < K; ++q) v[q] = bc[q];
  __syncthreads();
}

template <class T>
__global__ void __launch_bounds__(1024) k_coarse_global(MgLevel L, int singular, const T* __restrict__ b,
                                                        const double* mb, T* __restrict__ x, T* __restrict__ scr,
                                                        const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  __shared__ double sh[2 * 16];
  __shared__ double bc[2];
  const int N = int(L.n);
  T* r = scr;
  T* z = scr + N;
  T* p = scr + 2 * N;
  T* Ap = scr + 3 * N;
  const T bm = mb ? T(ldA(mb)) : T(0);
  const int tid = threadIdx.x, nt = blockDim.x;
  auto ijk = [&](int c, int& i, int& j, int& k) {
    i = c % L.nx;
    j = (c / L.nx) % L.ny;
    k = c / (L.nx * L.ny);
  };
  double s1[1] = {0.0};
  for (int c = tid; c < N; c += nt) {
    const bool fl = L.code[c] & kFluidBit;
    T v = b[c];
    if (mb && fl) v = v - bm;
    r[c] = v;
    x[c] = T(0);
    if (fl) s1[0] += double(v);
  }
  blockAll<1>(s1, sh, bc);
  if (singular && L.nf > 0) {
    const T m = T(s1[0] / double(L.nf));
    for (int c = tid; c < N; c += nt) r[c] = (L.code[c] & kFluidBit) ? r[c] - m : T(0);
  }
  __syncthreads();
  const XView<T> P = plainView<T>(p);
  const XView<T> Z = plainView<T>(r);
  auto diagOf = [&](int c, uint16_t code) {
    if (!(code & kFluidBit)) return T(1);
    int i, j, k;
    ijk(c, i, j, k);
    T dg;
    applyRow<T, true>(L, code, c, i, j, k, Z, dg);
    return dg > T(0) ? dg : T(1);
  };
  double s2[2] = {0.0, 0.0};
  for (int c = tid; c < N; c += nt) {
    const T zv = r[c] / diagOf(c, L.code[c]);
    z[c] = zv;
    p[c] = zv;
    s2[0] += double(r[c]) * double(zv);
    s2[1] += double(r[c]) * double(r[c]);
  }
  blockAll<2>(s2, sh, bc);
  double rz = s2[0];
  const double bn = sqrt(s2[1]) + 1e-300;
  for (int it = 0; it < 500; ++it) {
    double s[1] = {0.0};
    for (int c = tid; c < N; c += nt) {
      const uint16_t code = L.code[c];
      T v;
      if (code & kFluidBit) {
        int i, j, k;
        ijk(c, i, j, k);
        T dg;
        v = applyRow<T, false>(L, code, c, i, j, k, P, dg);
      } else {
        v = p[c];
      }
      Ap[c] = v;
      s[0] += double(p[c]) * double(v);
    }
    blockAll<1>(s, sh, bc);
    if (fabs(s[0]) < 1e-300) break;
    const double alpha = rz / s[0];
    double q2[2] = {0.0, 0.0};
    for (int c = tid; c < N; c += nt) {
      x[c] = x[c] + T(alpha) * p[c];
      const T rv = r[c] - T(alpha) * Ap[c];
      r[c] = rv;
      const T zv = rv / diagOf(c, L.code[c]);
      z[c] = zv;
      q2[0] += double(rv) * double(rv);
      q2[1] += double(rv) * double(zv);
    }
    blockAll<2>(q2, sh, bc);
    if (sqrt(q2[0]) <= 1e-6 * bn) break;
    const double beta = q2[1] / rz;
    rz = q2[1];
    for (int c = tid; c < N; c += nt) p[c] = z[c] + T(beta) * p[c];
    __syncthreads();
  }
  __syncthreads();
  if (singular && L.nf > 0) {
    double sx[1] = {0.0};
    for (int c = tid; c < N; c += nt)
      if (L.code[c] & kFluidBit) sx[0] += double(x[c]);
    blockAll<1>(sx, sh, bc);
    const T m = T(sx[0] / double(L.nf));
    for (int c = tid; c < N; c += nt)
      if (L.code[c] & kFluidBit) x[c] = x[c] - m;
  }
}

// Level-0 preconditioner output: z = x - pending mean (fluid), 0 elsewhere,
// with fluid partial sums for Minv's own projection (multigrid.cpp:283-289).
template <class T>
__global__ void __launch_bounds__(kBlock) k_finish(MgLevel L, const T* __restrict__ x, const double* mx,
                                                   double* __restrict__ z, int fold, RedBuf rb,
                                                   const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    const bool fl = L.code[q.c] & kFluidBit;
    T v = x[q.c];
    if (mx && fl) v = v - T(ldA(mx));
    const double zv = fl ? double(v) : 0.0;
    z[q.c] = zv;
    acc[0] = fl ? zv : 0.0;
  }
  if (fold) blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

template <class T>
__global__ void k_to_level0(int n, const double* __restrict__ r, T* __restrict__ b, const SolveCtl* ctl) {
  if (ctl && ldA(&ctl->done)) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) b[c] = T(r[c]);
}

// ---------------------------------------------------------------------------
// Plain operator pieces (kernel-level parity hooks)
// ---------------------------------------------------------------------------
__global__ void k_apply(MgLevel L, const double* __restrict__ x, double* __restrict__ out) {
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  if (!q.valid) return;
  const uint16_t code = L.code[q.c];
  if (!(code & kFluidBit)) {
    out[q.c] = x[q.c];
    return;
  }
  double dg;
  out[q.c] = applyRow<double, false>(L, code, q.c, q.i, q.j, q.k, plainView<double>(x), dg);
}

__global__ void k_diag(MgLevel L, double* __restrict__ out) {
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  if (!q.valid) return;
  const uint16_t code = L.code[q.c];
  if (!(code & kFluidBit)) {
    out[q.c] = 1.0;
    return;
  }
  double dg;
  applyRow<double, true>(L, code, q.c, q.i, q.j, q.k, plainView<double>(nullptr), dg);
  out[q.c] = dg > 0.0 ? dg : 1.0;
}

__global__ void k_prolong(MgLevel L, MgLevel Lc, const double* __restrict__ ec, double* __restrict__ xf) {
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  if (!q.valid || !(L.code[q.c] & kFluidBit)) return;
  xf[q.c] = xf[q.c] + ec[(q.i / Lc.fx) + Lc.nx * ((q.j / Lc.fy) + Lc.ny * (q.k / Lc.fz))];
}

// ---------------------------------------------------------------------------
// Flexible CG (MgPoisson::solve, multigrid.cpp:206-355) and Jacobi CG
// (CgPoisson::solve, poisson.cpp:136-252).  Scalars live in SolveCtl; each
// reducing kernel writes per-block partials and is followed by k_fin.
// ---------------------------------------------------------------------------
enum : int { S_MEAN_RHS = 0, S_MEAN_RN = 1, S_MEAN_X = 2 };

__global__ void __launch_bounds__(kBlock) k_fcg_init(MgLevel L, int maskSolid, const double* __restrict__ b,
                                                     double* __restrict__ rhs, double* __restrict__ x, RedBuf rb) {
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    const bool fl = L.code[q.c] & kFluidBit;
    const double v = (fl || !maskSolid) ? b[q.c] : 0.0;
    rhs[q.c] = v;
    x[q.c] = 0.0;
    acc[0] = fl ? v : 0.0;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// rhs <- rhs - mean (fluid) / 0 (solid) on singular systems; r = rhs; |rhs|^2.
__global__ void __launch_bounds__(kBlock) k_fcg_init2(MgLevel L, int singular, double* __restrict__ rhs,
                                                      double* __restrict__ r, const SolveCtl* ctl, RedBuf rb) {
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    double v = rhs[q.c];
    if (singular) {
      v = (L.code[q.c] & kFluidBit) ? v - ldA(&ctl->slot[S_MEAN_RHS]) : 0.0;
      rhs[q.c] = v;
    }
    r[q.c] = v;
    acc[0] = v * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// z <- z - pending mean (fluid); p = z; r.z
__global__ void __launch_bounds__(kBlock) k_fcg_start(MgLevel L, double* __restrict__ z, const double* mz,
                                                      const double* __restrict__ r, double* __restrict__ p,
                                                      const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    double v = z[q.c];
    if (mz && (L.code[q.c] & kFluidBit)) v = v - ldA(mz);
    z[q.c] = v;
    p[q.c] = v;
    acc[0] = r[q.c] * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

__global__ void __launch_bounds__(kBlock) k_fcg_apply(MgLevel L, const double* __restrict__ p,
                                                      double* __restrict__ Ap, const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    const uint16_t code = L.code[q.c];
    double v;
    if (code & kFluidBit) {
      double dg;
      v = applyRow<double, false>(L, code, q.c, q.i, q.j, q.k, plainView<double>(p), dg);
    } else {
      v = p[q.c];
    }
    Ap[q.c] = v;
    acc[0] = p[q.c] * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// x += alpha p; rn = r - alpha Ap; fluid sum of rn (singular) or |rn|^2.
__global__ void __launch_bounds__(kBlock) k_fcg_update(MgLevel L, int singular, double* __restrict__ x,
                                                       const double* __restrict__ p, const double* __restrict__ r,
                                                       const double* __restrict__ Ap, double* __restrict__ rn,
                                                       const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double a = ldA(&ctl->alpha);
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    x[q.c] = x[q.c] + a * p[q.c];
    const double v = r[q.c] - a * Ap[q.c];
    rn[q.c] = v;
    if (singular) acc[0] = (L.code[q.c] & kFluidBit) ? v : 0.0;
    else acc[0] = v * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

__global__ void __launch_bounds__(kBlock) k_fcg_norm(MgLevel L, double* __restrict__ rn, const SolveCtl* ctl,
                                                     RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    const double v = (L.code[q.c] & kFluidBit) ? rn[q.c] - ldA(&ctl->slot[S_MEAN_RN]) : 0.0;
    rn[q.c] = v;
    acc[0] = v * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// z <- z - pending mean; z.rn and z.r (flexible Polak-Ribiere beta,
// multigrid.cpp:328-339)
__global__ void __launch_bounds__(kBlock) k_fcg_dots(MgLevel L, double* __restrict__ z, const double* mz,
                                                     const double* __restrict__ rn, const double* __restrict__ r,
                                                     const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[2] = {0.0, 0.0};
  if (q.valid) {
    double v = z[q.c];
    if (mz && (L.code[q.c] & kFluidBit)) v = v - ldA(mz);
    z[q.c] = v;
    acc[0] = v * rn[q.c];
    acc[1] = v * r[q.c];
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, q.blk);
}

// k_finish + k_fcg_dots in one pass for systems without projections: z is the
// V-cycle output on fluid cells (converted to FP64), 0 elsewhere.
template <class T>
__global__ void __launch_bounds__(kBlock) k_fcg_dots_x(MgLevel L, const T* __restrict__ x, double* __restrict__ z,
                                                       const double* __restrict__ rn, const double* __restrict__ r,
                                                       const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[2] = {0.0, 0.0};
  if (q.valid) {
    const double v = (L.code[q.c] & kFluidBit) ? double(x[q.c]) : 0.0;
    z[q.c] = v;
    acc[0] = v * rn[q.c];
    acc[1] = v * r[q.c];
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, q.blk);
}

__global__ void __launch_bounds__(kBlock) k_fcg_pupdate(int n, double* __restrict__ p, const double* __restrict__ z,
                                                        double* __restrict__ r, const double* __restrict__ rn,
                                                        const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const double be = ldA(&ctl->beta);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  p[c] = z[c] + be * p[c];
  if (r) r[c] = rn[c];
}

__global__ void __launch_bounds__(kBlock) k_fluid_sum(MgLevel L, const double* __restrict__ x, RedBuf rb) {
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid && (L.code[q.c] & kFluidBit)) acc[0] = x[q.c];
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// Final projection of x and the true residual (multigrid.cpp:342-354).
__global__ void __launch_bounds__(kBlock) k_fcg_final(MgLevel L, int singular, const double* __restrict__ x,
                                                      const double* __restrict__ rhs, double* __restrict__ xo,
                                                      const SolveCtl* ctl, RedBuf rb) {
  const double m = singular ? ldA(&ctl->slot[S_MEAN_X]) : 0.0;
  const XView<double> X{x, m, singular != 0, nullptr, 1, 1, 0, 0, 0, 0.0, false};
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    const uint16_t code = L.code[q.c];
    double Ax, xv;
    if (code & kFluidBit) {
      double dg;
      Ax = applyRow<double, false>(L, code, q.c, q.i, q.j, q.k, X, dg);
      xv = singular ? x[q.c] - m : x[q.c];
    } else {
      xv = singular ? 0.0 : x[q.c];
      Ax = xv;
    }
    xo[q.c] = xv;
    const double e = rhs[q.c] - Ax;
    acc[0] = e * e;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// Jacobi-CG: z = r/diag and r.z (first: also p = z).
__global__ void __launch_bounds__(kBlock) k_cg_z(MgLevel L, const double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, int first, const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    const uint16_t code = L.code[q.c];
    double diag = 1.0;
    if (code & kFluidBit) {
      double dg;
      applyRow<double, true>(L, code, q.c, q.i, q.j, q.k, plainView<double>(nullptr), dg);
      diag = dg > 0.0 ? dg : 1.0;
    }
    const double zv = r[q.c] / diag;
    z[q.c] = zv;
    if (first) p[q.c] = zv;
    acc[0] = r[q.c] * zv;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

__global__ void __launch_bounds__(kBlock) k_cg_update(MgLevel L, double* __restrict__ x, const double* __restrict__ p,
                                                      double* __restrict__ r, const double* __restrict__ Ap,
                                                      const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double a = ldA(&ctl->alpha);
  const Cell q = cellOf(L.nx, L.ny, L.nz);
  double acc[1] = {0.0};
  if (q.valid) {
    x[q.c] = x[q.c] + a * p[q.c];
    const double v = r[q.c] - a * Ap[q.c];
    r[q.c] = v;
    acc[0] = v * v;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// One-block finalize: fold the partials of the preceding reducing kernel and
// apply the scalar update `op`.  Thread 0 reads the scalars coherently into a
// local copy, updates it, and writes it back.
__global__ void __launch_bounds__(kFinThreads) k_fin(RedBuf rb, FinArgs a, SolveCtl* ctl) {
  if (a.predicated && ldA(&ctl->done)) return;
  double t[2];
  if (a.K == 2) {
    foldPartials<2>(rb.partials, rb.stride, a.nparts, t);
  } else {
    double t1[1];
    foldPartials<1>(rb.partials, rb.stride, a.nparts, t1);
    t[0] = t1[0];
    t[1] = 0.0;
  }
  if (threadIdx.x != 0) return;
  double tol = ldA(&ctl->tol), bnorm = ldA(&ctl->bnorm), rz = ldA(&ctl->rz), pAp = ldA(&ctl->pAp);
  double alpha = ldA(&ctl->alpha), beta = ldA(&ctl->beta), rr = ldA(&ctl->rr), resid = ldA(&ctl->resid);
  int it = ldA(&ctl->it), done = ldA(&ctl->done), conv = ldA(&ctl->converged), broke = ldA(&ctl->broke);
  int early = ldA(&ctl->early);
  const double mean0 = a.nf > 0 ? t[0] / double(a.nf) : 0.0;
  const double mean1 = a.nf > 0 ? t[1] / double(a.nf) : 0.0;
  auto reset = [&] {
    it = 0;
    done = 0;
    conv = 0;
    broke = 0;
    early = 0;
  };
  switch (a.op) {
    case FIN_MEAN:
      stA(&ctl->slot[a.slot], mean0);
      return;
    case FIN_BNORM:  // multigrid.cpp:281, 292-298
      bnorm = sqrt(t[0]) + 1e-300;
      reset();
      if (a.earlyExit && sqrt(t[0]) <= tol * bnorm) {
        done = 1;
        conv = 1;
        early = 1;
        resid = sqrt(t[0]) / bnorm;
      }
      break;
    case FIN_RZ:
      rz = t[0];
      break;
    case FIN_ALPHA:
      pAp = t[0];
      if (fabs(t[0]) < 1e-300) {
        done = 1;
        broke = 1;
      } else {
        alpha = rz / t[0];
      }
      break;
    case FIN_CONV:
      rr = t[0];
      if (sqrt(t[0]) <= tol * bnorm) {
        it += 1;
        conv = 1;
        done = 1;
      }
      break;
    case FIN_BETA_FLEX:
      beta = (t[0] - t[1]) / rz;
      rz = t[0];
      it += 1;
      break;
    case FIN_BETA:
      beta = t[0] / rz;
      rz = t[0];
      it += 1;
      break;
    case FIN_RESID:
      if (early) return;  // returned before the loop: keep that report
      resid = sqrt(t[0]) / bnorm;
      conv = (conv || resid <= a.convTol) ? 1 : 0;
      break;
    case FIN_BNORM_MEAN:
      bnorm = sqrt(t[0]) + 1e-300;
      stA(&ctl->slot[a.slot], mean1);
      reset();
      break;
    case FIN_CONV_MEAN:
      stA(&ctl->slot[a.slot], mean1);
      rr = t[0];
      if (sqrt(t[0]) <= tol * bnorm) {
        it += 1;
        conv = 1;
        done = 1;
      }
      break;
  }
  stA(&ctl->bnorm, bnorm);
  stA(&ctl->rz, rz);
  stA(&ctl->pAp, pAp);
  stA(&ctl->alpha, alpha);
  stA(&ctl->beta, beta);
  stA(&ctl->rr, rr);
  stA(&ctl->resid, resid);
  stA(&ctl->it, it);
  stA(&ctl->converged, conv);
  stA(&ctl->broke, broke);
  stA(&ctl->early, early);
  stA(&ctl->done, done);
}

// Fresh solve scalars written by a kernel (not a host copy), so they reach
// memory through the same ordered, cache-coherent path as every later write.
__global__ void k_ctl_init(SolveCtl* ctl, double tol) {
  const int t = threadIdx.x;
  if (t < 64) stA(&ctl->slot[t], 0.0);
  if (t == 0) {
    stA(&ctl->tol, tol);
    for (double* f : {&ctl->bnorm, &ctl->rz, &ctl->pAp, &ctl->alpha, &ctl->beta, &ctl->rr, &ctl->resid}) stA(f, 0.0);
    for (int* f : {&ctl->it, &ctl->done, &ctl->converged, &ctl->broke, &ctl->early, &ctl->pad_}) stA(f, 0);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
void launch_ctl_init(SolveCtl* ctl, double tol, hipStream_t s) {
  k_ctl_init<<<1, 64, 0, s>>>(ctl, tol);
  NBL_LAUNCH_CHECK();
}

void launch_fin(RedBuf rb, const FinArgs& a, SolveCtl* ctl, hipStream_t s) {
  k_fin<<<1, kFinThreads, 0, s>>>(rb, a, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_sweep(const MgLevel& L, const T* b, const double* mb, const T* x, const double* mx,
                     bool zeroGuess, const T* ec, const double* mec, const MgLevel* Lc, T* xo, bool fold,
                     RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  const int ncx = Lc ? Lc->nx : 1, ncy = Lc ? Lc->ny : 1;
  const int sx = Lc ? Lc->fx - 1 : 0, sy = Lc ? Lc->fy - 1 : 0, sz = Lc ? Lc->fz - 1 : 0;
  const int f = fold ? 1 : 0;
  if (zeroGuess)
    k_sweep<T, SW_ZERO><<<g.grid, g.block, 0, s>>>(L, b, mb, x, mx, nullptr, nullptr, 1, 1, 0, 0, 0, xo, f, rb, ctl);
  else if (ec)
    k_sweep<T, SW_PROLONG><<<g.grid, g.block, 0, s>>>(L, b, mb, x, mx, ec, mec, ncx, ncy, sx, sy, sz, xo, f, rb, ctl);
  else
    k_sweep<T, SW_PLAIN><<<g.grid, g.block, 0, s>>>(L, b, mb, x, mx, nullptr, nullptr, 1, 1, 0, 0, 0, xo, f, rb, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_sweep_pair(const MgLevel& L, const T* b, T* xo, const SolveCtl* ctl, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_sweep<T, SW_PAIR><<<g.grid, g.block, 0, s>>>(L, b, nullptr, nullptr, nullptr, nullptr, nullptr, 1, 1, 0, 0, 0, xo,
                                                 0, RedBuf{nullptr, 0}, ctl);
  NBL_LAUNCH_CHECK();
}

Geo3 restrictGeo(const MgLevel& L, const MgLevel& Lc) {
  const int bz = Lc.fz;
  const int bx = L.nx >= 64 ? 64 : L.nx >= 32 ? 32 : 16;
  const int by = kBlock / (bx * bz);
  Geo3 g;
  g.block = dim3(bx, by, bz);
  g.grid = dim3((L.nx + bx - 1) / bx, (L.ny + by - 1) / by, (L.nz + bz - 1) / bz);
  g.blocks = int(g.grid.x * g.grid.y * g.grid.z);
  return g;
}

template <class T>
void launch_mg_resid_restrict(const MgLevel& L, const MgLevel& Lc, const T* b, const double* mb, const T* x,
                              const double* mx, T* rc, bool fold, RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  const Geo3 g = restrictGeo(L, Lc);
  k_resid_restrict<T, false><<<g.grid, g.block, 0, s>>>(L, Lc, b, mb, x, mx, rc, fold ? 1 : 0, rb, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_coarse(const MgLevel& L, bool singular, const T* b, const double* mb, T* x, T* scratch,
                      const SolveCtl* ctl, hipStream_t s) {
  if (L.n <= kCoarseMax)
    k_coarse_lds<T><<<1, kCoarseT, 0, s>>>(L, singular ? 1 : 0, b, mb, x, ctl);
  else
    k_coarse_global<T><<<1, 1024, 0, s>>>(L, singular ? 1 : 0, b, mb, x, scratch, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_finish(const MgLevel& L, const T* x, const double* mx, double* z, bool fold, RedBuf rb,
                      const SolveCtl* ctl, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_finish<T><<<g.grid, g.block, 0, s>>>(L, x, mx, z, fold ? 1 : 0, rb, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_to_level0(const MgLevel& L, const double* r, T* b, const SolveCtl* ctl, hipStream_t s) {
  k_to_level0<T><<<gridFor(L.n), kBlock, 0, s>>>(int(L.n), r, b, ctl);
  NBL_LAUNCH_CHECK();
}

void launch_gj_step(double* G, int n, int k, double* colk, double* rowk, int* fail, double tiny, hipStream_t s) {
  k_gj_copy<<<gridFor(2LL * n), kBlock, 0, s>>>(G, n, k, colk, rowk, fail, tiny);
  const dim3 grid((2 * n - k + kBlock - 1) / kBlock, n, 1);
  k_gj_update<<<grid, kBlock, 0, s>>>(G, n, k, colk, rowk);
  NBL_LAUNCH_CHECK();
}

void launch_inv_store(const double* G, int n, const uint16_t* code, long long nf, bool singular, double* scratch,
                      double* M, float* Mf, hipStream_t s) {
  if (singular) {
    k_inv_rowmean<<<n, kBlock, 0, s>>>(G, n, code, nf, scratch);
    k_inv_total<<<1, kBlock, 0, s>>>(scratch, n, code, nf, scratch + n);
  }
  const dim3 grid((n + kBlock - 1) / kBlock, n, 1);
  k_inv_store<<<grid, kBlock, 0, s>>>(G, n, code, singular ? 1 : 0, scratch, scratch + n, M, Mf);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_coarse_direct(const MgLevel& L, const double* M, const float* Mf, const T* b, const double* mb, T* x,
                             const SolveCtl* ctl, hipStream_t s) {
  const int blocks = int((L.n + kDirectRows - 1) / kDirectRows);
  if constexpr (sizeof(T) == sizeof(float))
    k_coarse_direct<T, float><<<blocks, kBlock, 0, s>>>(L, Mf, b, mb, x, ctl);
  else
    k_coarse_direct<T, double><<<blocks, kBlock, 0, s>>>(L, M, b, mb, x, ctl);
  NBL_LAUNCH_CHECK();
}

#define NBL_INST(T)                                                                                          \
  template void launch_mg_sweep<T>(const MgLevel&, const T*, const double*, const T*, const double*, bool,    \
                                   const T*, const double*, const MgLevel*, T*, bool, RedBuf, const SolveCtl*, \
                                   hipStream_t);                                                             \
  template void launch_mg_resid_restrict<T>(const MgLevel&, const MgLevel&, const T*, const double*,         \
                                            const T*, const double*, T*, bool, RedBuf, const SolveCtl*,      \
                                            hipStream_t);                                                    \
  template void launch_mg_coarse<T>(const MgLevel&, bool, const T*, const double*, T*, T*,                   \
                                    const SolveCtl*, hipStream_t);                                           \
  template void launch_mg_finish<T>(const MgLevel&, const T*, const double*, double*, bool, RedBuf,          \
                                    const SolveCtl*, hipStream_t);                                           \
  template void launch_to_level0<T>(const MgLevel&, const double*, T*, const SolveCtl*, hipStream_t);        \
  template void launch_mg_sweep_pair<T>(const MgLevel&, const T*, T*, const SolveCtl*, hipStream_t);        \
  template void launch_mg_coarse_direct<T>(const MgLevel&, const double*, const float*, const T*,            \
                                           const double*, T*, const SolveCtl*, hipStream_t);
NBL_INST(double)
NBL_INST(float)
#undef NBL_INST

void launch_apply(const MgLevel& L, const double* x, double* out, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_apply<<<g.grid, g.block, 0, s>>>(L, x, out);
  NBL_LAUNCH_CHECK();
}
void launch_diag(const MgLevel& L, double* out, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_diag<<<g.grid, g.block, 0, s>>>(L, out);
  NBL_LAUNCH_CHECK();
}
void launch_restrict_plain(const MgLevel& L, const MgLevel& Lc, const double* rf, double* rc, hipStream_t s) {
  // rf - A*0 == rf exactly, so the fused kernel with a zero x restricts rf itself
  const Geo3 g = restrictGeo(L, Lc);
  k_resid_restrict<double, true><<<g.grid, g.block, 0, s>>>(L, Lc, rf, nullptr, nullptr, nullptr, rc, 0,
                                                            RedBuf{nullptr, 0}, nullptr);
  NBL_LAUNCH_CHECK();
}
void launch_prolong_plain(const MgLevel& L, const MgLevel& Lc, const double* ec, double* xf, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_prolong<<<g.grid, g.block, 0, s>>>(L, Lc, ec, xf);
  NBL_LAUNCH_CHECK();
}

void launch_fcg_init(const MgLevel& L, bool maskSolid, const double* b, double* rhs, double* x, RedBuf rb,
                     hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_init<<<g.grid, g.block, 0, s>>>(L, maskSolid ? 1 : 0, b, rhs, x, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_init2(const MgLevel& L, bool singular, double* rhs, double* r, const SolveCtl* ctl, RedBuf rb,
                      hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_init2<<<g.grid, g.block, 0, s>>>(L, singular ? 1 : 0, rhs, r, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_start(const MgLevel& L, double* z, const double* mz, const double* r, double* p,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_start<<<g.grid, g.block, 0, s>>>(L, z, mz, r, p, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_apply(const MgLevel& L, const double* p, double* Ap, const SolveCtl* ctl, RedBuf rb,
                      hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_apply<<<g.grid, g.block, 0, s>>>(L, p, Ap, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_update(const MgLevel& L, bool singular, double* x, const double* p, const double* r,
                       const double* Ap, double* rn, const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_update<<<g.grid, g.block, 0, s>>>(L, singular ? 1 : 0, x, p, r, Ap, rn, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_norm(const MgLevel& L, double* rn, const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_norm<<<g.grid, g.block, 0, s>>>(L, rn, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_dots(const MgLevel& L, double* z, const double* mz, const double* rn, const double* r,
                     const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_dots<<<g.grid, g.block, 0, s>>>(L, z, mz, rn, r, ctl, rb);
  NBL_LAUNCH_CHECK();
}
template <class T>
void launch_fcg_dots_x(const MgLevel& L, const T* x, double* z, const double* rn, const double* r,
                       const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_dots_x<T><<<g.grid, g.block, 0, s>>>(L, x, z, rn, r, ctl, rb);
  NBL_LAUNCH_CHECK();
}
template void launch_fcg_dots_x<double>(const MgLevel&, const double*, double*, const double*, const double*,
                                        const SolveCtl*, RedBuf, hipStream_t);
template void launch_fcg_dots_x<float>(const MgLevel&, const float*, double*, const double*, const double*,
                                       const SolveCtl*, RedBuf, hipStream_t);
void launch_fcg_pupdate(const MgLevel& L, double* p, const double* z, double* r, const double* rn,
                        const SolveCtl* ctl, hipStream_t s) {
  k_fcg_pupdate<<<gridFor(L.n), kBlock, 0, s>>>(int(L.n), p, z, r, rn, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_fluid_sum(const MgLevel& L, const double* x, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fluid_sum<<<g.grid, g.block, 0, s>>>(L, x, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_final(const MgLevel& L, bool singular, const double* x, const double* rhs, double* xo,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_fcg_final<<<g.grid, g.block, 0, s>>>(L, singular ? 1 : 0, x, rhs, xo, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_cg_z(const MgLevel& L, const double* r, double* z, bool first, double* p, const SolveCtl* ctl,
                 RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_cg_z<<<g.grid, g.block, 0, s>>>(L, r, z, p, first ? 1 : 0, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_cg_update(const MgLevel& L, double* x, const double* p, double* r, const double* Ap,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_cg_update<<<g.grid, g.block, 0, s>>>(L, x, p, r, Ap, ctl, rb);
  NBL_LAUNCH_CHECK();
}

}  // namespace nbl
