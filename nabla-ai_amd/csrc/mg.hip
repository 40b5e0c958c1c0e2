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
#include <algorithm>
#include <stdexcept>
#include <type_traits>

#include "mgkern.hpp"

namespace nbl {

Geo3 geo3n(int nx, int ny, int nz, int cells, int slab, int nb) {
  const int tx = (nx + cells - 1) / cells;
  // the row width (64, 32, 16 or 8 threads; 256/width rows of j) that leaves
  // the fewest idle threads in the padded tx x ny rectangle, the widest of
  // equals (a width that divides the row exactly wastes nothing along i)
  int bx = 8;
  double best = -1.0;
  for (int w : {64, 32, 16, 8}) {
    if (w > 8 && w > 2 * tx) continue;
    const int h = kBlock / w;
    const double eff = double(tx) / double((tx + w - 1) / w * w) * double(ny) / double((ny + h - 1) / h * h);
    if (eff > best + 1e-12) {
      best = eff;
      bx = w;
    }
  }
  const int by = kBlock / bx;
  Geo3 g;
  g.block = dim3(bx, by, 1);
  const unsigned gx = (tx + bx - 1) / bx, gy = (ny + by - 1) / by;
  if (nb > 1 && slab == 1) throw std::logic_error("a batch runs its levels in plain order or bands of rows");
  const unsigned gz = unsigned(nz) * unsigned(std::max(nb, 1));
  g.grid = slab == 1 ? dim3(8 * gx, gy, (nz + 7) / 8) : slab == 2 ? dim3(8 * gx, (gy + 7) / 8, gz) : dim3(gx, gy, gz);
  g.blocks = int(g.grid.x * g.grid.y * g.grid.z);
  return g;
}
Geo3 geo3(int nx, int ny, int nz) { return geo3n(nx, ny, nz, 1); }
Geo3 levelGeo(const MgLevel& L) { return geo3n(L.nx, L.ny, L.nz, L.vec, L.slab, L.nb); }
int slabFor(int nx, int ny, int nz) {
  if ((long long)nx * ny * nz < (8LL << 20)) return nz >= 8 ? 1 : 0;
  return 2;
}

namespace {
template <class T, int MODE, int NC>
__global__ void __launch_bounds__(kBlock) k_sweep(MgLevel L, const T* __restrict__ b, MeanRef mb,
                                                  const T* __restrict__ x, MeanRef mx, const T* ec,
                                                  MeanRef mec, int ncx, int ncy, int sx, int sy, int sz, int ecN,
                                                  T* __restrict__ xo, int fold, RedBuf rb, SolveCtl* ctl,
                                                  long long* stamps, InFold conv) {
  const CellN<NC> q = cellOfN<NC>(L.nx, L.ny, L.nz, L.slab, L.nb);
  if (L.nb > 1) {  // this block's problem: its SolveCtl, mean slots and coarse correction
    ctl = ctlAt(ctl, q.p);
    mb = mb.at(q.p);
    mx = mx.at(q.p);
    mec = mec.at(q.p);
    if (ec) ec += size_t(q.p) * size_t(ecN);
  }
  const bool t0 = threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0;
  if (stamps && t0) stamps[2 * q.blk] = wall_clock64();
  // Only a sweep from zero carries the convergence test (the host passes it
  // to no other mode); it has no pending means, so its cell loads go out
  // before the test's fold and the two memory round trips overlap.
  uint16_t code[NC];
  T xc[NC], xn[NC][6], bv[NC];
  double mv[3] = {0.0, 0.0, 0.0};
  if constexpr (MODE == SW_ZERO) {
    if (!conv.n && stopOrMeans<3>(ctl, {mb, mx, mec}, mv)) return;
  } else {
    if (stopOrMeans<3>(ctl, {mb, mx, mec}, mv)) return;
  }
  if (q.valid) {
    ldCodes<NC>(L.code, q.c, code);
    if constexpr (MODE == SW_PROLONG) {
      const XView<T, true> X{x, T(mv[1]), mx.on(), ec, ncx, ncy, sx, sy, sz, T(mv[2]), mec.on()};
      gatherN<T, NC>(L, X, q.c, q.i, q.j, q.k, xc, xn);
    } else if constexpr (MODE == SW_PLAIN) {
      const XView<T> X{x, T(mv[1]), mx.on(), nullptr, 1, 1, 0, 0, 0, T(0), false};
      gatherN<T, NC>(L, X, q.c, q.i, q.j, q.k, xc, xn);
    }
    ldN<NC>(b, q.c, bv);
  }
  if constexpr (MODE == SW_ZERO) {
    if (conv.n && inlineConv(conv, ctl)) return;
  }
  double acc[1] = {0.0};
  T out[NC];
  if (q.valid) {
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      if (!(code[e] & kFluidBit)) {
        out[e] = T(0);
        continue;
      }
      T dg, Ax, x0;
      if constexpr (MODE == SW_ZERO) {
        const T zn[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        Ax = applyDecode<T, true>(L, code[e], T(0), zn, dg);
        x0 = T(0);
      } else {
        Ax = applyDecode<T, false>(L, code[e], xc[e], xn[e], dg);
        x0 = xc[e];
      }
      const T diag = dg > T(0) ? dg : T(1);
      T bb = bv[e];
      if (mb.on()) bb = bb - T(mv[0]);
      const T xnew = x0 + T(0.8) * (bb - Ax) / diag;
      out[e] = xnew;
      acc[0] += double(xnew);
    }
  }
  // stores after the fold's barriers, which would otherwise wait for them
  if (fold) blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) stN<NC>(xo, q.c, out);
  if (stamps) {  // uniform across the grid: end stamp once every wave of the block is done
    __syncthreads();
    if (t0) stamps[2 * q.blk + 1] = wall_clock64();
  }
}

// Residual r = b - A x on fine fluid cells restricted to the coarse grid as
// the mean over fluid children (multigrid.cpp:116-154).  NC fine cells per
// thread along i (NC = 2 only with an x coarsening factor of 2: a thread then
// holds both x-children of one coarse cell); a block covers whole coarse
// cells (blockDim.z = fz, even x/y extents), stages its residuals in LDS, and
// one thread per coarse cell sums the children in the reference's k, j, i
// order.  ZEROX: x == 0 (restriction of b itself).  x1c (optional): the
// coarse level's first sweep from zero, which is pointwise (k_sweep<SW_ZERO>'s
// expression on the restricted value), so the coarse level starts at its
// second sweep.
template <class T, bool ZEROX, int NC>
__global__ void __launch_bounds__(kBlock) k_resid_restrict(MgLevel L, MgLevel Lc, const T* __restrict__ b,
                                                           MeanRef mb, const T* __restrict__ x, MeanRef mx,
                                                           T* __restrict__ rc, int fold, RedBuf rb, SolveCtl* ctl,
                                                           T* __restrict__ x1c, InFold conv) {
  __shared__ T rs[kBlock * NC];
  __shared__ unsigned char fls[kBlock * NC];
  const Blk3 B = blk3(L.slab);
  const int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
  const int bx = wgSizeX(), by = wgSizeY(), bz = wgSizeZ();
  // a batch's problems follow each other along z, each in its own blocks
  int bzl = B.z, p = 0;
  if (L.nb > 1) {
    const int gzp = (L.nz + bz - 1) / bz;
    p = bzl / gzp;
    bzl -= p * gzp;
    ctl = ctlAt(ctl, p);
    mb = mb.at(p);
    mx = mx.at(p);
  }
  const int i = NC * (B.x * bx + tx), j = B.y * by + ty, k = bzl * bz + tz;
  const int tid = tx + bx * (ty + by * tz);
  const bool in = i < L.nx && j < L.ny && k < L.nz;
  const int c = i + L.nx * (j + L.ny * (k + L.nz * p));
  uint16_t code[NC];
  T xc[NC], xn[NC][6], bv[NC];
  // the coarse cell this thread sums (one thread per coarse cell, the first
  // of its children), and its code, loaded with the fine stencil so the sum
  // after the barrier waits for no further memory round trip
  const int fxT = NC == 2 ? 1 : Lc.fx;  // x-children held by one thread: 2 (NC = 2) or 1
  int Cst = -1;
  uint16_t cc = 0;
  if (tz == 0 && tx % fxT == 0 && ty % Lc.fy == 0) {
    const int I = i / Lc.fx, J = j / Lc.fy, K = k / Lc.fz;
    if (I < Lc.nx && J < Lc.ny && K < Lc.nz) {
      Cst = I + Lc.nx * (J + Lc.ny * (K + Lc.nz * p));
      cc = Lc.code[Cst];
    }
  }
  // a restriction that carries the convergence test: as k_sweep
  double mv[2] = {0.0, 0.0};
  if (!conv.n && stopOrMeans<2>(ctl, {mb, mx}, mv)) return;
  if (in) {
    const XView<T> X{x, T(mv[1]), mx.on(), nullptr, 1, 1, 0, 0, 0, T(0), false};
    ldCodes<NC>(L.code, c, code);
    if constexpr (!ZEROX) gatherN<T, NC>(L, X, c, i, j, k, xc, xn);
    ldN<NC>(b, c, bv);
  }
  if (conv.n && inlineConv(conv, ctl)) return;
  T r[NC];
  bool fl[NC];
#pragma unroll
  for (int e = 0; e < NC; ++e) {
    r[e] = T(0);
    fl[e] = false;
  }
  if (in) {
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      if (!(code[e] & kFluidBit)) continue;
      fl[e] = true;
      T dg, Ax;
      if constexpr (ZEROX) {
        const T zn[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        Ax = applyDecode<T, true>(L, code[e], T(0), zn, dg);
      } else {
        Ax = applyDecode<T, false>(L, code[e], xc[e], xn[e], dg);
      }
      T bb = bv[e];
      if (mb.on()) bb = bb - T(mv[0]);
      r[e] = bb - Ax;
    }
  }
#pragma unroll
  for (int e = 0; e < NC; ++e) {
    rs[NC * tid + e] = r[e];
    fls[NC * tid + e] = fl[e] ? 1 : 0;
  }
  __syncthreads();
  double acc[1] = {0.0};
  T rcv = T(0), x1v = T(0);  // the coarse cell's values, stored after the fold
  if (Cst >= 0) {
    const int I = i / Lc.fx, J = j / Lc.fy, K = k / Lc.fz;
    const int C = Cst;
    {
      if (cc & kFluidBit) {
        T s = T(0);
        int cnt = 0;
        for (int dk = 0; dk < Lc.fz; ++dk)
          for (int dj = 0; dj < Lc.fy; ++dj)
            for (int di = 0; di < Lc.fx; ++di) {
              const int t = NC == 2 ? 2 * (tx + bx * ((ty + dj) + by * dk)) + di
                                    : (tx + di) + bx * ((ty + dj) + by * dk);
              if (fls[t]) {
                s += rs[t];
                ++cnt;
              }
            }
        const T rv = s / T(cnt);
        rcv = rv;
        if (x1c) x1v = firstSweep<T>(Lc, cc, C, I, J, K, rv);
        acc[0] = double(rv);
      }
    }
  }
  if (fold) blockToPartials<1>(acc, rb.partials, rb.stride, B.lin);
  if (Cst >= 0) {
    rc[Cst] = rcv;
    if (x1c) x1c[Cst] = x1v;
  }
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

// x = Minv (b - mb on fluid cells): ROWS rows per block, each thread
// striding the columns; fixed-order block fold.
template <class T, class MT, int ROWS>
__global__ void __launch_bounds__(kBlock) k_coarse_direct(MgLevel L, const MT* __restrict__ M,
                                                          const T* __restrict__ b, MeanRef mb,
                                                          T* __restrict__ x, const SolveCtl* ctl) {
  const int n = int(L.n);
  int blk = blockIdx.x;
  if (L.nb > 1) {  // ceil(n / ROWS) blocks per problem of a batch; one inverse for all
    const int bpp = (n + ROWS - 1) / ROWS;
    const int p = blk / bpp;
    blk -= p * bpp;
    ctl = ctlAt(ctl, p);
    mb = mb.at(p);
    b += size_t(p) * size_t(n);
    x += size_t(p) * size_t(n);
  }
  double mv[1];
  if (stopOrMeans<1>(ctl, {mb}, mv)) return;
  const int r0 = blk * ROWS;
  const double m = mv[0];
  double acc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) acc[r] = 0.0;
  // the loads of four column strides are issued before their products are
  // accumulated (in column order), one memory round trip per four strides
  constexpr int CB = 4;
  const int nt = wgSizeX();
  for (int j0 = int(threadIdx.x); j0 < n; j0 += CB * nt) {
    double bv[CB], mv2[CB][ROWS];
#pragma unroll
    for (int u = 0; u < CB; ++u) {
      const int j = j0 + u * nt;
      bv[u] = 0.0;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) mv2[u][r] = 0.0;
      if (j < n) {
        T bj = b[j];
        if (mb.on() && (L.code[j] & kFluidBit)) bj = bj - T(m);
        bv[u] = double(bj);
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
          if (r0 + r < n) mv2[u][r] = double(M[size_t(r0 + r) * n + j]);
      }
    }
#pragma unroll
    for (int u = 0; u < CB; ++u)
      if (j0 + u * nt < n)
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
          if (r0 + r < n) acc[r] = fma(mv2[u][r], bv[u], acc[r]);
  }
  __shared__ double sh[ROWS * 16];
  blockFold<ROWS, 0u>(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
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
template <class T, int NC>
__global__ void __launch_bounds__(kBlock) k_finish(MgLevel L, const T* __restrict__ x, MeanRef mx,
                                                   double* __restrict__ z, int fold, RedBuf rb,
                                                   const SolveCtl* ctl) {
  const CellN<NC> q = cellOfN<NC>(L.nx, L.ny, L.nz, L.slab, L.nb);
  if (L.nb > 1) {
    ctl = ctlAt(ctl, q.p);
    mx = mx.at(q.p);
  }
  double mv[1];
  if (stopOrMeans<1>(ctl, {mx}, mv)) return;
  double acc[1] = {0.0};
  double zv[NC];
  if (q.valid) {
    uint16_t code[NC];
    T xv[NC];
    ldCodes<NC>(L.code, q.c, code);
    ldN<NC>(x, q.c, xv);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      const bool fl = code[e] & kFluidBit;
      T v = xv[e];
      if (mx.on() && fl) v = v - T(mv[0]);
      zv[e] = fl ? double(v) : 0.0;
      acc[0] += zv[e];
    }
  }
  if (fold) blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) stN<NC>(z, q.c, zv);  // after the fold's barriers
}

// n cells per problem; a batch's nb problems one after another (problem c / n).
template <class T>
__global__ void k_to_level0(int n, int nb, const double* __restrict__ r, T* __restrict__ b, const SolveCtl* ctl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n * nb) return;
  if (ctl && ldA(&ctlAt(ctl, nb > 1 ? c / n : 0)->done)) return;
  b[c] = T(r[c]);
}

// ---------------------------------------------------------------------------
// The reference's multigrid and flexible-CG loops, one kernel per loop
// (multigrid.cpp:81-194, 268-355; poisson.cpp:102-134): the unfused FP64
// form, a measurement baseline for what the fused kernels above save.  Every
// kernel is one loop of the reference over all cells (grid-stride), does
// nothing once the solve has stopped, and reductions write per-block partials.
// ---------------------------------------------------------------------------
#define NBL_LOOP(c, n) for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < (n); \
                            c += (long long)gridDim.x * blockDim.x)

// CgPoisson::applyA: out = A x, identity rows on solid cells
__global__ void k_loop_apply(MgLevel L, const double* __restrict__ x, double* __restrict__ out, const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const long long sz = (long long)L.nx * L.ny;
  NBL_LOOP(c, L.n) {
    const int k = int(c / sz), j = int((c - k * sz) / L.nx), i = int(c - k * sz - (long long)j * L.nx);
    const uint16_t code = L.code[c];
    if (!(code & kFluidBit)) {
      out[c] = x[c];
      continue;
    }
    double dg;
    out[c] = applyRow<double, false>(L, code, int(c), i, j, k, plainView<double>(x), dg);
  }
}
// smooth's update loop: x = fluid ? x + 0.8 (b - Ax) / diag : 0
__global__ void k_loop_jacobi(MgLevel L, const double* __restrict__ b, const double* __restrict__ Ax,
                              const double* __restrict__ diag, double* __restrict__ x, const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  NBL_LOOP(c, L.n) x[c] = (L.code[c] & kFluidBit) ? x[c] + 0.8 * (b[c] - Ax[c]) / diag[c] : 0.0;
}
// removeMean, first loop: fluid sum, solid cells set to 0
__global__ void k_loop_msum(MgLevel L, double* __restrict__ x, RedBuf rb, const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  double acc[1] = {0.0};
  NBL_LOOP(c, L.n) {
    if (L.code[c] & kFluidBit) acc[0] += x[c];
    else x[c] = 0.0;
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, blockIdx.x);
}
// removeMean, second loop: fluid cells minus the mean (a SolveCtl slot)
__global__ void k_loop_msub(MgLevel L, double* __restrict__ x, const double* mean, const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const double m = ldA(mean);
  NBL_LOOP(c, L.n) if (L.code[c] & kFluidBit) x[c] -= m;
}
// residual's loop: r = fluid ? b - r : 0 (r holds A x)
__global__ void k_loop_resid(MgLevel L, const double* __restrict__ b, double* __restrict__ r, const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  NBL_LOOP(c, L.n) r[c] = (L.code[c] & kFluidBit) ? b[c] - r[c] : 0.0;
}
// restrictResidual: each coarse cell the mean of its fluid children, summed
// in the fine loop's k, j, i order
__global__ void k_loop_restrict(MgLevel L, MgLevel Lc, const double* __restrict__ rf, double* __restrict__ rc,
                                const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const long long szc = (long long)Lc.nx * Lc.ny;
  NBL_LOOP(C, Lc.n) {
    const int K = int(C / szc), J = int((C - K * szc) / Lc.nx), I = int(C - K * szc - (long long)J * Lc.nx);
    double sum = 0.0;
    int cnt = 0;
    for (int k = K * Lc.fz; k < min(K * Lc.fz + Lc.fz, L.nz); ++k)
      for (int j = J * Lc.fy; j < min(J * Lc.fy + Lc.fy, L.ny); ++j)
        for (int i = I * Lc.fx; i < min(I * Lc.fx + Lc.fx, L.nx); ++i) {
          const long long f = i + (long long)L.nx * (j + (long long)L.ny * k);
          if (L.code[f] & kFluidBit) {
            sum += rf[f];
            ++cnt;
          }
        }
    rc[C] = cnt > 0 ? sum / double(cnt) : 0.0;
  }
}
// prolongAdd: fluid fine cells add their parent's correction
__global__ void k_loop_prolong(MgLevel L, MgLevel Lc, const double* __restrict__ ec, double* __restrict__ xf,
                               const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const long long sz = (long long)L.nx * L.ny;
  NBL_LOOP(c, L.n) {
    if (!(L.code[c] & kFluidBit)) continue;
    const int k = int(c / sz), j = int((c - k * sz) / L.nx), i = int(c - k * sz - (long long)j * L.nx);
    xf[c] += ec[(i / Lc.fx) + (long long)Lc.nx * ((j / Lc.fy) + (long long)Lc.ny * (k / Lc.fz))];
  }
}
// dot(a, b); K = 2: z.rn and z.r in one loop (the flexible beta's loop)
__global__ void k_loop_dot(long long n, const double* __restrict__ a, const double* __restrict__ b,
                           const double* __restrict__ b2, RedBuf rb, const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  double acc[2] = {0.0, 0.0};
  NBL_LOOP(c, n) {
    acc[0] += a[c] * b[c];
    if (b2) acc[1] += a[c] * b2[c];
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, blockIdx.x);
}
// x += alpha p; rnew = r - alpha Ap
__global__ void k_loop_update(long long n, double* __restrict__ x, const double* __restrict__ p,
                              const double* __restrict__ r, const double* __restrict__ Ap,
                              double* __restrict__ rn, const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const double a = ldA(&ctl->alpha);
  NBL_LOOP(c, n) {
    x[c] += a * p[c];
    rn[c] = r[c] - a * Ap[c];
  }
}
// p = znew + beta p
__global__ void k_loop_pupdate(long long n, double* __restrict__ p, const double* __restrict__ z,
                               const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  const double be = ldA(&ctl->beta);
  NBL_LOOP(c, n) p[c] = z[c] + be * p[c];
}
// dst = src (r = rnew), dst = 0 (zz.assign)
__global__ void k_loop_copy(long long n, double* __restrict__ dst, const double* __restrict__ src,
                            const SolveCtl* ctl) {
  if (ldA(&ctl->done)) return;
  NBL_LOOP(c, n) dst[c] = src ? src[c] : 0.0;
}
#undef NBL_LOOP

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
// reducing kernel writes per-block partials and is followed by k_fin.  All
// run NC cells per thread along i (the level's vec) on levelGeo's launch.
// ---------------------------------------------------------------------------
enum : int { S_MEAN_RHS = 0, S_MEAN_RN = 1, S_MEAN_X = 2 };

// The thread's cells and their codes; in a batch the block's problem (q.p).
#define NBL_CELLS(NC)                                               \
  const CellN<NC> q = cellOfN<NC>(L.nx, L.ny, L.nz, L.slab, L.nb);  \
  uint16_t code[NC];                                                \
  if (q.valid) ldCodes<NC>(L.code, q.c, code)

// The first V-cycle's level-0 inputs beside r: its FP32 right-hand side
// (b0, the mixed form) and its first sweep from a zero guess (x1), each
// optional, as k_to_level0 and k_sweep<SW_ZERO> write them.
template <class T, int NC>
__device__ __forceinline__ void firstCycleInputs(const MgLevel& L, const CellN<NC>& q, const uint16_t (&code)[NC],
                                                 const double (&v)[NC], float* b0, T* x1) {
  if (b0) {
    float f[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) f[e] = float(v[e]);
    stN<NC>(b0, q.c, f);
  }
  if (x1) {
    T fv[NC];
#pragma unroll
    for (int e = 0; e < NC; ++e) fv[e] = firstSweep<T>(L, code[e], q.c + e, q.i + e, q.j, q.k, T(v[e]));
    stN<NC>(x1, q.c, fv);
  }
}

// rhs = b (0 on solid cells where maskSolid), x = 0, partials of rhs's fluid
// sum.  With r (systems without projection): also r = rhs, the partials then
// |rhs|^2, and the first cycle's inputs (b0, x1) beside them.
template <class T, int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_init(MgLevel L, int maskSolid, const double* __restrict__ b,
                                                     double* __restrict__ rhs, double* __restrict__ x,
                                                     double* __restrict__ r, float* __restrict__ b0,
                                                     T* __restrict__ x1, RedBuf rb) {
  NBL_CELLS(NC);
  double acc[1] = {0.0};
  if (q.valid) {
    double bv[NC], v[NC], z[NC];
    ldN<NC>(b, q.c, bv);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      const bool fl = code[e] & kFluidBit;
      v[e] = (fl || !maskSolid) ? bv[e] : 0.0;
      z[e] = 0.0;
      acc[0] += r ? v[e] * v[e] : (fl ? v[e] : 0.0);
    }
    stN<NC>(rhs, q.c, v);
    stN<NC>(x, q.c, z);
    if (r) {
      stN<NC>(r, q.c, v);
      firstCycleInputs<T, NC>(L, q, code, v, b0, x1);
    }
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// rhs <- rhs - mean (fluid) / 0 (solid) on singular systems; r = rhs; |rhs|^2;
// the first cycle's inputs (b0, x1) beside them.
template <class T, int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_init2(MgLevel L, int singular, double* __restrict__ rhs,
                                                      double* __restrict__ r, const SolveCtl* ctl, RedBuf rb,
                                                      float* __restrict__ b0, T* __restrict__ x1) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  double acc[1] = {0.0};
  if (q.valid) {
    double v[NC];
    ldN<NC>(rhs, q.c, v);
    if (singular) {
      const double m = ldA(&ctl->slot[S_MEAN_RHS]);
#pragma unroll
      for (int e = 0; e < NC; ++e) v[e] = (code[e] & kFluidBit) ? v[e] - m : 0.0;
      stN<NC>(rhs, q.c, v);
    }
    stN<NC>(r, q.c, v);
#pragma unroll
    for (int e = 0; e < NC; ++e) acc[0] += v[e] * v[e];
    firstCycleInputs<T, NC>(L, q, code, v, b0, x1);
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// z <- z - pending mean (fluid); p = z; r.z
template <int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_start(MgLevel L, double* __restrict__ z, MeanRef mz,
                                                      const double* __restrict__ r, double* __restrict__ p,
                                                      const SolveCtl* ctl, RedBuf rb) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  mz = mz.at(q.p);
  double mv[1];
  if (stopOrMeans<1>(ctl, {mz}, mv)) return;
  double acc[1] = {0.0};
  if (q.valid) {
    double v[NC], rv[NC];
    ldN<NC>(z, q.c, v);
    ldN<NC>(r, q.c, rv);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      if (mz.on() && (code[e] & kFluidBit)) v[e] = v[e] - mv[0];
      acc[0] += rv[e] * v[e];
    }
    stN<NC>(z, q.c, v);
    stN<NC>(p, q.c, v);
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// Ap (identity rows on solid cells) and p.Ap, p read through the view P.
template <int NC, class PV>
__device__ __forceinline__ void applyN(const MgLevel& L, const CellN<NC>& q, const uint16_t (&code)[NC],
                                       const PV& P, double (&pv)[NC], double (&Ap)[NC]) {
  double xn[NC][6];
  gatherN<double, NC>(L, P, q.c, q.i, q.j, q.k, pv, xn);
#pragma unroll
  for (int e = 0; e < NC; ++e) {
    if (code[e] & kFluidBit) {
      double dg;
      Ap[e] = applyDecode<double, false>(L, code[e], pv[e], xn[e], dg);
    } else {
      Ap[e] = pv[e];
    }
  }
}

template <int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_apply(MgLevel L, const double* __restrict__ p,
                                                      double* __restrict__ Ap, const SolveCtl* ctl, RedBuf rb) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  if (ldA(&ctl->done)) return;
  double acc[1] = {0.0};
  double pv[NC], av[NC];
  if (q.valid) {
    applyN<NC>(L, q, code, plainView<double>(p), pv, av);
#pragma unroll
    for (int e = 0; e < NC; ++e) acc[0] += pv[e] * av[e];
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) stN<NC>(Ap, q.c, av);  // after the fold's barriers, which would wait for it
}

// p_new = z + beta p_old (multigrid.cpp:336-338) fused into the next
// iteration's A p and p.Ap: every stencil point re-forms z + beta p_old with
// the same expression, so A p_new is bit-identical to applying A to the
// stored p_new.  p_old and p_new are different buffers.
struct PNewView {
  const double* __restrict__ z;
  const double* __restrict__ p;
  double be;
  __device__ __forceinline__ double at(int c, int, int, int) const { return z[c] + be * p[c]; }
  template <int NC>
  __device__ __forceinline__ void pack(int c, int, int, int, double (&v)[NC]) const {
    double zv[NC], pv[NC];
    ldN<NC>(z, c, zv);
    ldN<NC>(p, c, pv);
#pragma unroll
    for (int e = 0; e < NC; ++e) v[e] = zv[e] + be * pv[e];
  }
};

// The same with z's fluid mean m still pending (singular systems whose last
// level-0 sweep folded z's fluid sum beside the dots): p_new = (z - m) + beta
// p_old, the projection's and the update's expressions in their order.  A
// non-fluid neighbour's value is never read by the operator row; a non-fluid
// centre's p_new and A p are zero, as projected z would give.
struct PNewMeanView {
  const double* __restrict__ z;
  const double* __restrict__ p;
  double be, m;
  __device__ __forceinline__ double at(int c, int, int, int) const { return (z[c] - m) + be * p[c]; }
  template <int NC>
  __device__ __forceinline__ void pack(int c, int, int, int, double (&v)[NC]) const {
    double zv[NC], pv[NC];
    ldN<NC>(z, c, zv);
    ldN<NC>(p, c, pv);
#pragma unroll
    for (int e = 0; e < NC; ++e) v[e] = (zv[e] - m) + be * pv[e];
  }
};

// MZ: z's mean is pending; with inline folds beta's partials carry a third
// row, z's fluid sum (nf = the fluid cell count), otherwise beta and the mean
// are read from ctl and the slot mz.
template <int NC, bool MZ>
__global__ void __launch_bounds__(kBlock) k_fcg_apply_pu(MgLevel L, const double* __restrict__ z,
                                                         const double* __restrict__ pOld, double* __restrict__ pNew,
                                                         double* __restrict__ Ap, SolveCtl* ctl, RedBuf rb,
                                                         InFold bf, double nf, MeanRef mz) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);  // (batches fold no partials inline: bf.n == 0)
  mz = mz.at(q.p);
  double be, m = 0.0;
  if (MZ && bf.n) {
    if (!inlineBetaMean(bf, ctl, nf, be, m)) return;
  } else if (MZ) {
    if (ldA(&ctl->done)) return;
    be = ldA(&ctl->beta);
    m = ldA(mz.slot);
  } else if (bf.n) {
    if (!inlineBeta(bf, ctl, be)) return;
  } else {
    // `done` is not written by this kernel: the per-thread read is uniform
    if (ldA(&ctl->done)) return;
    be = ldA(&ctl->beta);
  }
  double acc[1] = {0.0};
  double pv[NC], av[NC];
  if (q.valid) {
    if constexpr (MZ) {
      applyN<NC>(L, q, code, PNewMeanView{z, pOld, be, m}, pv, av);
#pragma unroll
      for (int e = 0; e < NC; ++e)
        if (!(code[e] & kFluidBit)) pv[e] = av[e] = 0.0;
    } else {
      applyN<NC>(L, q, code, PNewView{z, pOld, be}, pv, av);
    }
#pragma unroll
    for (int e = 0; e < NC; ++e) acc[0] += pv[e] * av[e];
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) {
    stN<NC>(pNew, q.c, pv);
    stN<NC>(Ap, q.c, av);
  }
}

// x += alpha p; rn = r - alpha Ap; fluid sum of rn (singular) or |rn|^2.
template <int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_update(MgLevel L, int singular, double* __restrict__ x,
                                                       const double* __restrict__ p, const double* __restrict__ r,
                                                       const double* __restrict__ Ap, double* __restrict__ rn,
                                                       SolveCtl* ctl, RedBuf rb, InFold af) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  double xv[NC], pv[NC], rv[NC], av[NC];
  auto load = [&] {
    if (q.valid) {
      ldN<NC>(x, q.c, xv);
      ldN<NC>(p, q.c, pv);
      ldN<NC>(r, q.c, rv);
      ldN<NC>(Ap, q.c, av);
    }
  };
  double a;
  if (af.n) {
    load();  // before the fold, so the two memory round trips overlap
    if (!inlineAlpha(af, ctl, a)) return;
  } else {
    if (ldA(&ctl->done)) return;
    a = ldA(&ctl->alpha);
    load();
  }
  double acc[1] = {0.0};
  if (q.valid) {
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      xv[e] = xv[e] + a * pv[e];
      rv[e] = rv[e] - a * av[e];
      if (singular) acc[0] += (code[e] & kFluidBit) ? rv[e] : 0.0;
      else acc[0] += rv[e] * rv[e];
    }
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) {  // after the fold's barriers, which would wait for the stores
    stN<NC>(x, q.c, xv);
    stN<NC>(rn, q.c, rv);
  }
}

// k_fcg_update without r_new's projection (a non-singular system's r_new
// has none; a singular one's stays mean-free to rounding) plus the first
// smoother sweep of the next V-cycle from a zero guess, which is pointwise (x1 = 0.8 rn / diag on
// fluid cells, k_sweep<SW_ZERO>'s expression), and for an FP32 V-cycle the
// level-0 right-hand side float(rn); partials of |rn|^2.
template <class T, int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_update_first(MgLevel L, double* __restrict__ x,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ Ap, double* __restrict__ rn,
                                                             T* __restrict__ x1, T* __restrict__ b0,
                                                             SolveCtl* ctl, RedBuf rb, InFold af) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  double xv[NC], pv[NC], rv[NC], av[NC];
  auto load = [&] {
    if (q.valid) {
      ldN<NC>(x, q.c, xv);
      ldN<NC>(p, q.c, pv);
      ldN<NC>(r, q.c, rv);
      ldN<NC>(Ap, q.c, av);
    }
  };
  double a;
  if (af.n) {
    load();  // before the fold, so the two memory round trips overlap
    if (!inlineAlpha(af, ctl, a)) return;
  } else {
    if (ldA(&ctl->done)) return;
    a = ldA(&ctl->alpha);
    load();
  }
  double acc[1] = {0.0};
  T bv[NC], fv[NC];
  if (q.valid) {
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      xv[e] = xv[e] + a * pv[e];
      rv[e] = rv[e] - a * av[e];
      bv[e] = T(rv[e]);
      fv[e] = firstSweep<T>(L, code[e], q.c + e, q.i + e, q.j, q.k, bv[e]);
      acc[0] += rv[e] * rv[e];
    }
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) {  // after the fold's barriers, which would wait for the stores
    stN<NC>(x, q.c, xv);
    stN<NC>(rn, q.c, rv);
    if (b0) stN<NC>(b0, q.c, bv);
    stN<NC>(x1, q.c, fv);
  }
}

// r_new <- r_new - mean (fluid) / 0; |r_new|^2; the FP32 V-cycle's right-hand
// side (b0) and the level-0 first sweep from a zero guess in the cycle's type
// (x1) beside it, each optional.
template <class T, int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_norm(MgLevel L, double* __restrict__ rn, MeanRef mrn,
                                                     const SolveCtl* ctl, RedBuf rb, float* __restrict__ b0,
                                                     T* __restrict__ x1) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  mrn = mrn.at(q.p);
  double mv[1];
  if (stopOrMeans<1>(ctl, {mrn}, mv)) return;
  double acc[1] = {0.0};
  double v[NC];
  if (q.valid) {
    const double m = mv[0];
    ldN<NC>(rn, q.c, v);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      v[e] = (code[e] & kFluidBit) ? v[e] - m : 0.0;
      acc[0] += v[e] * v[e];
    }
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) {  // after the fold's barriers, which would wait for the stores
    stN<NC>(rn, q.c, v);
    if (b0) {  // the FP32 V-cycle's right-hand side, as k_to_level0 writes it
      float f[NC];
#pragma unroll
      for (int e = 0; e < NC; ++e) f[e] = float(v[e]);
      stN<NC>(b0, q.c, f);
    }
    if (x1) {  // the first sweep from zero on that right-hand side, as k_sweep<T, SW_ZERO> writes it
      T fv[NC];
#pragma unroll
      for (int e = 0; e < NC; ++e) fv[e] = firstSweep<T>(L, code[e], q.c + e, q.i + e, q.j, q.k, T(v[e]));
      stN<NC>(x1, q.c, fv);
    }
  }
}

// z <- z - pending mean; z.rn and z.r (flexible Polak-Ribiere beta,
// multigrid.cpp:328-339)
template <int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_dots(MgLevel L, double* __restrict__ z, MeanRef mz,
                                                     const double* __restrict__ rn, const double* __restrict__ r,
                                                     SolveCtl* ctl, RedBuf rb, int countIt) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  mz = mz.at(q.p);
  double mv[1];
  if (stopOrMeans<1>(ctl, {mz}, mv)) return;
  if (countIt && problemLead(q, L.nz, L.nb) && linearTid() == 0) stA(&ctl->it, ldA(&ctl->it) + 1);
  double acc[2] = {0.0, 0.0};
  double v[NC];
  if (q.valid) {
    double a[NC], c[NC];
    ldN<NC>(z, q.c, v);
    ldN<NC>(rn, q.c, a);
    ldN<NC>(r, q.c, c);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      if (mz.on() && (code[e] & kFluidBit)) v[e] = v[e] - mv[0];
      acc[0] += v[e] * a[e];
      acc[1] += v[e] * c[e];
    }
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) stN<NC>(z, q.c, v);  // after the fold's barriers
}

// k_finish + k_fcg_dots in one pass: z is the V-cycle output on fluid cells
// (converted to FP64, minus its pending fluid mean on singular systems, as
// k_finish and k_fcg_dots form it), 0 elsewhere.
template <class T, int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_dots_x(MgLevel L, const T* __restrict__ x, MeanRef mx,
                                                       double* __restrict__ z, const double* __restrict__ rn,
                                                       const double* __restrict__ r, SolveCtl* ctl, RedBuf rb,
                                                       int countIt) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  mx = mx.at(q.p);
  double mv[1];
  if (stopOrMeans<1>(ctl, {mx}, mv)) return;
  if (countIt && problemLead(q, L.nz, L.nb) && linearTid() == 0) stA(&ctl->it, ldA(&ctl->it) + 1);
  double acc[2] = {0.0, 0.0};
  double v[NC];
  if (q.valid) {
    T xv[NC];
    double a[NC], c[NC];
    ldN<NC>(x, q.c, xv);
    ldN<NC>(rn, q.c, a);
    ldN<NC>(r, q.c, c);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      v[e] = (code[e] & kFluidBit) ? double(xv[e]) : 0.0;
      if (mx.on() && (code[e] & kFluidBit)) v[e] = v[e] - mv[0];
      acc[0] += v[e] * a[e];
      acc[1] += v[e] * c[e];
    }
  }
  blockToPartials<2>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) stN<NC>(z, q.c, v);  // after the fold's barriers
}

// The last level-0 sweep of a V-cycle with the preconditioner output and the
// flexible-CG dots folded in: z = the sweep's result (FP64) on fluid cells, 0
// elsewhere; partials of z.rn and z.r.  Same values as k_sweep<SW_PLAIN>
// followed by k_fcg_dots_x.  ZS (singular systems): a third row of partials
// holds z's fluid sum, and z's mean is left pending for the next A p; r_new
// and r are mean-free to rounding, so z.rn and z.r equal the projected z's.
template <class T, int NC, bool ZS>
__global__ void __launch_bounds__(kBlock) k_sweep_dots(MgLevel L, const T* __restrict__ b, const T* __restrict__ x,
                                                       double* __restrict__ z, const double* __restrict__ rn,
                                                       const double* __restrict__ r, RedBuf rb, SolveCtl* ctl,
                                                       int countIt) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  if (ldA(&ctl->done)) return;
  if (countIt && problemLead(q, L.nz, L.nb) && linearTid() == 0) stA(&ctl->it, ldA(&ctl->it) + 1);
  constexpr int K = ZS ? 3 : 2;
  double acc[K] = {};
  double v[NC];
  if (q.valid) {
    T xc[NC], xn[NC][6], bv[NC];
    double a[NC], c[NC];
    gatherN<T, NC>(L, plainView<T>(x), q.c, q.i, q.j, q.k, xc, xn);
    ldN<NC>(b, q.c, bv);
    ldN<NC>(rn, q.c, a);
    ldN<NC>(r, q.c, c);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      v[e] = 0.0;
      if (code[e] & kFluidBit) {
        T dg;
        const T Ax = applyDecode<T, false>(L, code[e], xc[e], xn[e], dg);
        const T diag = dg > T(0) ? dg : T(1);
        v[e] = double(xc[e] + T(0.8) * (bv[e] - Ax) / diag);
      }
      acc[0] += v[e] * a[e];
      acc[1] += v[e] * c[e];
      if constexpr (ZS) acc[K - 1] += v[e];
    }
  }
  blockToPartials<K>(acc, rb.partials, rb.stride, q.blk);
  if (q.valid) stN<NC>(z, q.c, v);  // after the fold's barriers
}

template <int NC>
__global__ void __launch_bounds__(kBlock) k_fluid_sum(MgLevel L, const double* __restrict__ x, RedBuf rb) {
  NBL_CELLS(NC);
  double acc[1] = {0.0};
  if (q.valid) {
    double v[NC];
    ldN<NC>(x, q.c, v);
#pragma unroll
    for (int e = 0; e < NC; ++e)
      if (code[e] & kFluidBit) acc[0] += v[e];
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// Final projection of x and the true residual (multigrid.cpp:342-354).
template <int NC>
__global__ void __launch_bounds__(kBlock) k_fcg_final(MgLevel L, int singular, const double* __restrict__ x,
                                                      const double* __restrict__ rhs, double* __restrict__ xo,
                                                      const SolveCtl* ctl, RedBuf rb) {
  NBL_CELLS(NC);
  ctl = ctlAt(ctl, q.p);
  const double m = singular ? ldA(&ctl->slot[S_MEAN_X]) : 0.0;
  const XView<double> X{x, m, singular != 0, nullptr, 1, 1, 0, 0, 0, 0.0, false};
  double acc[1] = {0.0};
  if (q.valid) {
    double xc[NC], xn[NC][6], raw[NC], hv[NC], out[NC];
    gatherN<double, NC>(L, X, q.c, q.i, q.j, q.k, xc, xn);
    ldN<NC>(x, q.c, raw);
    ldN<NC>(rhs, q.c, hv);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      double Ax;
      if (code[e] & kFluidBit) {
        double dg;
        Ax = applyDecode<double, false>(L, code[e], xc[e], xn[e], dg);
        out[e] = singular ? raw[e] - m : raw[e];
      } else {
        out[e] = singular ? 0.0 : raw[e];
        Ax = out[e];
      }
      const double er = hv[e] - Ax;
      acc[0] += er * er;
    }
    stN<NC>(xo, q.c, out);
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

// Jacobi-CG: z = r/diag and r.z (first: also p = z).
template <int NC>
__global__ void __launch_bounds__(kBlock) k_cg_z(MgLevel L, const double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, int first, const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  NBL_CELLS(NC);
  double acc[1] = {0.0};
  if (q.valid) {
    double rv[NC], zv[NC];
    ldN<NC>(r, q.c, rv);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      double diag = 1.0;
      if (code[e] & kFluidBit) {
        double dg;
        const double zn[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        applyDecode<double, true>(L, code[e], 0.0, zn, dg);
        diag = dg > 0.0 ? dg : 1.0;
      }
      zv[e] = rv[e] / diag;
      acc[0] += rv[e] * zv[e];
    }
    stN<NC>(z, q.c, zv);
    if (first) stN<NC>(p, q.c, zv);
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

template <int NC>
__global__ void __launch_bounds__(kBlock) k_cg_update(MgLevel L, double* __restrict__ x, const double* __restrict__ p,
                                                      double* __restrict__ r, const double* __restrict__ Ap,
                                                      const SolveCtl* ctl, RedBuf rb) {
  if (ldA(&ctl->done)) return;
  const double a = ldA(&ctl->alpha);
  NBL_CELLS(NC);
  double acc[1] = {0.0};
  if (q.valid) {
    double xv[NC], pv[NC], rv[NC], av[NC];
    ldN<NC>(x, q.c, xv);
    ldN<NC>(p, q.c, pv);
    ldN<NC>(r, q.c, rv);
    ldN<NC>(Ap, q.c, av);
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      xv[e] = xv[e] + a * pv[e];
      rv[e] = rv[e] - a * av[e];
      acc[0] += rv[e] * rv[e];
    }
    stN<NC>(x, q.c, xv);
    stN<NC>(r, q.c, rv);
  }
  blockToPartials<1>(acc, rb.partials, rb.stride, q.blk);
}

#undef NBL_CELLS

// One-block finalize: fold the partials of the preceding reducing kernel and
// apply the scalar update `op`.  Thread 0 reads the scalars coherently into a
// local copy, updates it, and writes it back.
__global__ void __launch_bounds__(kFinThreads) k_fin(RedBuf rb, FinArgs a, SolveCtl* ctl) {
  // block p finalizes problem p of a batch from its producer blocks' partials
  ctl += blockIdx.x;
  const double* parts = rb.partials + size_t(blockIdx.x) * size_t(a.nparts);
  if (a.predicated && ldA(&ctl->done)) return;
  double t[3] = {0.0, 0.0, 0.0};
  if (a.K == 3) {
    foldPartials<3>(parts, rb.stride, a.nparts, t);
  } else if (a.K == 2) {
    double t2[2];
    foldPartials<2>(parts, rb.stride, a.nparts, t2);
    t[0] = t2[0];
    t[1] = t2[1];
  } else {
    double t1[1];
    foldPartials<1>(parts, rb.stride, a.nparts, t1);
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
      stA(&ctl->rzq[0], rz);
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
    case FIN_BETA_FLEX_MEAN:
      stA(&ctl->slot[a.slot], a.nf > 0 ? t[2] / double(a.nf) : 0.0);
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
    case FIN_BNORM_R0:  // warm start: the stopping test's |b| and the first residual
      bnorm = sqrt(t[0]) + 1e-300;
      reset();
      if (a.earlyExit && sqrt(t[1]) <= tol * bnorm) {
        done = 1;
        conv = 1;
        early = 1;
        resid = sqrt(t[1]) / bnorm;
      }
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
  ctl += blockIdx.x;  // one block per problem of a batch
  const int t = threadIdx.x;
  if (t < 64) stA(&ctl->slot[t], 0.0);
  if (t == 0) {
    stA(&ctl->tol, tol);
    for (double* f : {&ctl->bnorm, &ctl->rz, &ctl->pAp, &ctl->alpha, &ctl->beta, &ctl->rr, &ctl->resid, &ctl->rzq[0],
                      &ctl->rzq[1]})
      stA(f, 0.0);
    for (int* f : {&ctl->it, &ctl->done, &ctl->converged, &ctl->broke, &ctl->early, &ctl->pad_}) stA(f, 0);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
void launch_ctl_init(SolveCtl* ctl, double tol, hipStream_t s, int nb) {
  k_ctl_init<<<std::max(nb, 1), 64, 0, s>>>(ctl, tol);
  NBL_LAUNCH_CHECK();
}

void launch_fin(RedBuf rb, const FinArgs& a, SolveCtl* ctl, hipStream_t s, int nb) {
  k_fin<<<std::max(nb, 1), kFinThreads, 0, s>>>(rb, a, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_sweep(const MgLevel& L, const T* b, MeanRef mb, const T* x, MeanRef mx,
                     bool zeroGuess, const T* ec, MeanRef mec, const MgLevel* Lc, T* xo, bool fold,
                     RedBuf rb, SolveCtl* ctl, hipStream_t s, long long* stamps, InFold conv) {
  const int ncx = Lc ? Lc->nx : 1, ncy = Lc ? Lc->ny : 1;
  const int sx = Lc ? Lc->fx - 1 : 0, sy = Lc ? Lc->fy - 1 : 0, sz = Lc ? Lc->fz - 1 : 0;
  const int ecN = Lc ? int(Lc->n) : 0;
  const int f = fold ? 1 : 0;
  if (conv.n && !zeroGuess) throw std::logic_error("mg sweep: only a sweep from zero carries the convergence test");
  withCells(L, [&](auto nc, const Geo3& g) {
    constexpr int NC = decltype(nc)::value;
    if (zeroGuess)
      k_sweep<T, SW_ZERO, NC><<<g.grid, g.block, 0, s>>>(L, b, mb, x, mx, nullptr, MeanRef{}, 1, 1, 0, 0, 0, 0, xo,
                                                         f, rb, ctl, stamps, conv);
    else if (ec)
      k_sweep<T, SW_PROLONG, NC><<<g.grid, g.block, 0, s>>>(L, b, mb, x, mx, ec, mec, ncx, ncy, sx, sy, sz, ecN, xo,
                                                            f, rb, ctl, stamps, conv);
    else
      k_sweep<T, SW_PLAIN, NC><<<g.grid, g.block, 0, s>>>(L, b, mb, x, mx, nullptr, MeanRef{}, 1, 1, 0, 0, 0, 0, xo,
                                                          f, rb, ctl, stamps, conv);
  });
  NBL_LAUNCH_CHECK();
}

// Fine cells per thread of the restriction from L to Lc.
static int restrictCells(const MgLevel& L, const MgLevel& Lc) { return (L.vec >= 2 && Lc.fx == 2) ? 2 : 1; }

Geo3 restrictGeo(const MgLevel& L, const MgLevel& Lc) {
  const int nc = restrictCells(L, Lc);
  const int tx = L.nx / nc;
  const int bz = Lc.fz;
  // widest row of threads that divides the row exactly, else the widest
  int bx = tx >= 64 ? 64 : tx >= 32 ? 32 : 16;
  for (int w : {64, 32, 16})
    if (w <= tx && tx % w == 0) {
      bx = w;
      break;
    }
  const int by = kBlock / (bx * bz);
  Geo3 g;
  g.block = dim3(bx, by, bz);
  const unsigned gx = (tx + bx - 1) / bx, gy = (L.ny + by - 1) / by, gz = (L.nz + bz - 1) / bz;
  if (L.nb > 1 && L.slab == 1) throw std::logic_error("a batch runs its levels in plain order or bands of rows");
  const unsigned gzb = gz * unsigned(std::max(L.nb, 1));  // a batch's problems follow each other along z
  g.grid = L.slab == 1 ? dim3(8 * gx, gy, (gz + 7) / 8) : L.slab == 2 ? dim3(8 * gx, (gy + 7) / 8, gzb) : dim3(gx, gy, gzb);
  g.blocks = int(g.grid.x * g.grid.y * g.grid.z);
  return g;
}

template <class T>
void launch_mg_resid_restrict(const MgLevel& L, const MgLevel& Lc, const T* b, MeanRef mb, const T* x,
                              MeanRef mx, T* rc, bool fold, RedBuf rb, SolveCtl* ctl, hipStream_t s, T* x1c,
                              InFold conv) {
  const Geo3 g = restrictGeo(L, Lc);
  if (restrictCells(L, Lc) == 2)
    k_resid_restrict<T, false, 2><<<g.grid, g.block, 0, s>>>(L, Lc, b, mb, x, mx, rc, fold ? 1 : 0, rb, ctl, x1c,
                                                             conv);
  else
    k_resid_restrict<T, false, 1><<<g.grid, g.block, 0, s>>>(L, Lc, b, mb, x, mx, rc, fold ? 1 : 0, rb, ctl, x1c,
                                                             conv);
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
void launch_mg_finish(const MgLevel& L, const T* x, MeanRef mx, double* z, bool fold, RedBuf rb,
                      const SolveCtl* ctl, hipStream_t s) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_finish<T, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, x, mx, z, fold ? 1 : 0, rb, ctl);
  });
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_to_level0(const MgLevel& L, const double* r, T* b, const SolveCtl* ctl, hipStream_t s) {
  const int nb = std::max(L.nb, 1);
  k_to_level0<T><<<gridFor(L.n * nb), kBlock, 0, s>>>(int(L.n), nb, r, b, ctl);
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
void launch_mg_coarse_direct(const MgLevel& L, const double* M, const float* Mf, const T* b, MeanRef mb, T* x,
                             const SolveCtl* ctl, hipStream_t s) {
  // two rows per block: measured fastest of 1, 2, 4, 8 rows per block and one
  // wavefront per row on the 1536-cell cube-channel level (5.1 us vs 6.7 us
  // for 4 rows, back-to-back)
  constexpr int kRows = 2;
  const int blocks = int((L.n + kRows - 1) / kRows) * std::max(L.nb, 1);
  if constexpr (sizeof(T) == sizeof(float))
    k_coarse_direct<T, float, kRows><<<blocks, kBlock, 0, s>>>(L, Mf, b, mb, x, ctl);
  else
    k_coarse_direct<T, double, kRows><<<blocks, kBlock, 0, s>>>(L, M, b, mb, x, ctl);
  NBL_LAUNCH_CHECK();
}

#define NBL_INST(T)                                                                                          \
  template void launch_mg_sweep<T>(const MgLevel&, const T*, MeanRef, const T*, MeanRef, bool, const T*,     \
                                   MeanRef, const MgLevel*, T*, bool, RedBuf, SolveCtl*, hipStream_t,        \
                                   long long*, InFold);                                                      \
  template void launch_mg_resid_restrict<T>(const MgLevel&, const MgLevel&, const T*, MeanRef, const T*,     \
                                            MeanRef, T*, bool, RedBuf, SolveCtl*, hipStream_t, T*, InFold);  \
  template void launch_mg_coarse<T>(const MgLevel&, bool, const T*, const double*, T*, T*,                   \
                                    const SolveCtl*, hipStream_t);                                           \
  template void launch_mg_finish<T>(const MgLevel&, const T*, MeanRef, double*, bool, RedBuf,               \
                                    const SolveCtl*, hipStream_t);                                           \
  template void launch_to_level0<T>(const MgLevel&, const double*, T*, const SolveCtl*, hipStream_t);        \
  template void launch_mg_coarse_direct<T>(const MgLevel&, const double*, const float*, const T*, MeanRef,   \
                                           T*, const SolveCtl*, hipStream_t);
NBL_INST(double)
NBL_INST(float)
#undef NBL_INST

int loopBlocks(long long n) { return gridFor(n, kBlock, kLoopBlocks); }
void launch_loop_apply(const MgLevel& L, const double* x, double* out, const SolveCtl* ctl, hipStream_t s) {
  k_loop_apply<<<loopBlocks(L.n), kBlock, 0, s>>>(L, x, out, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_jacobi(const MgLevel& L, const double* b, const double* Ax, const double* diag, double* x,
                        const SolveCtl* ctl, hipStream_t s) {
  k_loop_jacobi<<<loopBlocks(L.n), kBlock, 0, s>>>(L, b, Ax, diag, x, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_msum(const MgLevel& L, double* x, RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  k_loop_msum<<<loopBlocks(L.n), kBlock, 0, s>>>(L, x, rb, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_msub(const MgLevel& L, double* x, const double* mean, const SolveCtl* ctl, hipStream_t s) {
  k_loop_msub<<<loopBlocks(L.n), kBlock, 0, s>>>(L, x, mean, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_resid(const MgLevel& L, const double* b, double* r, const SolveCtl* ctl, hipStream_t s) {
  k_loop_resid<<<loopBlocks(L.n), kBlock, 0, s>>>(L, b, r, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_restrict(const MgLevel& L, const MgLevel& Lc, const double* rf, double* rc, const SolveCtl* ctl,
                          hipStream_t s) {
  k_loop_restrict<<<loopBlocks(Lc.n), kBlock, 0, s>>>(L, Lc, rf, rc, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_prolong(const MgLevel& L, const MgLevel& Lc, const double* ec, double* xf, const SolveCtl* ctl,
                         hipStream_t s) {
  k_loop_prolong<<<loopBlocks(L.n), kBlock, 0, s>>>(L, Lc, ec, xf, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_dot(long long n, const double* a, const double* b, const double* b2, RedBuf rb,
                     const SolveCtl* ctl, hipStream_t s) {
  k_loop_dot<<<loopBlocks(n), kBlock, 0, s>>>(n, a, b, b2, rb, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_update(long long n, double* x, const double* p, const double* r, const double* Ap, double* rn,
                        const SolveCtl* ctl, hipStream_t s) {
  k_loop_update<<<loopBlocks(n), kBlock, 0, s>>>(n, x, p, r, Ap, rn, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_pupdate(long long n, double* p, const double* z, const SolveCtl* ctl, hipStream_t s) {
  k_loop_pupdate<<<loopBlocks(n), kBlock, 0, s>>>(n, p, z, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_loop_copy(long long n, double* dst, const double* src, const SolveCtl* ctl, hipStream_t s) {
  k_loop_copy<<<loopBlocks(n), kBlock, 0, s>>>(n, dst, src, ctl);
  NBL_LAUNCH_CHECK();
}

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
  if (restrictCells(L, Lc) == 2)
    k_resid_restrict<double, true, 2><<<g.grid, g.block, 0, s>>>(L, Lc, rf, MeanRef{}, nullptr, MeanRef{}, rc, 0,
                                                                 RedBuf{nullptr, 0}, nullptr, nullptr, InFold{});
  else
    k_resid_restrict<double, true, 1><<<g.grid, g.block, 0, s>>>(L, Lc, rf, MeanRef{}, nullptr, MeanRef{}, rc, 0,
                                                                 RedBuf{nullptr, 0}, nullptr, nullptr, InFold{});
  NBL_LAUNCH_CHECK();
}
void launch_prolong_plain(const MgLevel& L, const MgLevel& Lc, const double* ec, double* xf, hipStream_t s) {
  const Geo3 g = geo3(L.nx, L.ny, L.nz);
  k_prolong<<<g.grid, g.block, 0, s>>>(L, Lc, ec, xf);
  NBL_LAUNCH_CHECK();
}

void launch_fcg_init(const MgLevel& L, bool maskSolid, const double* b, double* rhs, double* x, RedBuf rb,
                     hipStream_t s, double* r, float* b0, float* x1f, double* x1d) {
  if (x1f && x1d) throw std::logic_error("fcg_init: one first-sweep output");
  if (!r && (b0 || x1f || x1d)) throw std::logic_error("fcg_init: first-cycle inputs need r");
  withCells(L, [&](auto nc, const Geo3& g) {
    if (x1f)
      k_fcg_init<float, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, maskSolid ? 1 : 0, b, rhs, x, r, b0, x1f,
                                                                       rb);
    else
      k_fcg_init<double, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, maskSolid ? 1 : 0, b, rhs, x, r, b0,
                                                                        x1d, rb);
  });
  NBL_LAUNCH_CHECK();
}
void launch_fcg_init2(const MgLevel& L, bool singular, double* rhs, double* r, const SolveCtl* ctl, RedBuf rb,
                      hipStream_t s, float* b0, float* x1f, double* x1d) {
  if (x1f && x1d) throw std::logic_error("fcg_init2: one first-sweep output");
  withCells(L, [&](auto nc, const Geo3& g) {
    if (x1f)
      k_fcg_init2<float, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, singular ? 1 : 0, rhs, r, ctl, rb, b0,
                                                                        x1f);
    else
      k_fcg_init2<double, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, singular ? 1 : 0, rhs, r, ctl, rb, b0,
                                                                         x1d);
  });
  NBL_LAUNCH_CHECK();
}
void launch_fcg_start(const MgLevel& L, double* z, MeanRef mz, const double* r, double* p,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fcg_start<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, z, mz, r, p, ctl, rb);
  });
  NBL_LAUNCH_CHECK();
}
void launch_fcg_apply(const MgLevel& L, const double* p, double* Ap, const SolveCtl* ctl, RedBuf rb,
                      hipStream_t s) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fcg_apply<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, p, Ap, ctl, rb);
  });
  NBL_LAUNCH_CHECK();
}
void launch_fcg_update(const MgLevel& L, bool singular, double* x, const double* p, const double* r,
                       const double* Ap, double* rn, SolveCtl* ctl, RedBuf rb, hipStream_t s, InFold alpha) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fcg_update<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, singular ? 1 : 0, x, p, r, Ap, rn, ctl, rb,
                                                                 alpha);
  });
  NBL_LAUNCH_CHECK();
}
template <class T>
void launch_fcg_update_first(const MgLevel& L, double* x, const double* p, const double* r, const double* Ap,
                             double* rn, T* x1, T* b0, SolveCtl* ctl, RedBuf rb, hipStream_t s, InFold alpha) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fcg_update_first<T, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, x, p, r, Ap, rn, x1, b0, ctl, rb, alpha);
  });
  NBL_LAUNCH_CHECK();
}
template void launch_fcg_update_first<double>(const MgLevel&, double*, const double*, const double*, const double*,
                                              double*, double*, double*, SolveCtl*, RedBuf, hipStream_t, InFold);
template void launch_fcg_update_first<float>(const MgLevel&, double*, const double*, const double*, const double*,
                                             double*, float*, float*, SolveCtl*, RedBuf, hipStream_t, InFold);
void launch_fcg_apply_pu(const MgLevel& L, const double* z, const double* pOld, double* pNew, double* Ap,
                         SolveCtl* ctl, RedBuf rb, hipStream_t s, InFold beta, double zMeanNf, const double* zMeanSlot) {
  if (zMeanNf > 0.0 && !beta.n && !zMeanSlot)
    throw std::logic_error("apply_pu: a pending z mean needs inline folds or its slot");
  withCells(L, [&](auto nc, const Geo3& g) {
    if (zMeanNf > 0.0)
      k_fcg_apply_pu<decltype(nc)::value, true><<<g.grid, g.block, 0, s>>>(L, z, pOld, pNew, Ap, ctl, rb, beta,
                                                                          zMeanNf, slotMean(zMeanSlot));
    else
      k_fcg_apply_pu<decltype(nc)::value, false><<<g.grid, g.block, 0, s>>>(L, z, pOld, pNew, Ap, ctl, rb, beta,
                                                                           0.0, MeanRef{});
  });
  NBL_LAUNCH_CHECK();
}
template <class T>
void launch_mg_sweep_dots(const MgLevel& L, const T* b, const T* x, double* z, const double* rn, const double* r,
                          RedBuf rb, SolveCtl* ctl, hipStream_t s, bool countIt, bool zSum) {
  withCells(L, [&](auto nc, const Geo3& g) {
    if (zSum)
      k_sweep_dots<T, decltype(nc)::value, true><<<g.grid, g.block, 0, s>>>(L, b, x, z, rn, r, rb, ctl,
                                                                           countIt ? 1 : 0);
    else
      k_sweep_dots<T, decltype(nc)::value, false><<<g.grid, g.block, 0, s>>>(L, b, x, z, rn, r, rb, ctl,
                                                                            countIt ? 1 : 0);
  });
  NBL_LAUNCH_CHECK();
}
template void launch_mg_sweep_dots<double>(const MgLevel&, const double*, const double*, double*, const double*,
                                           const double*, RedBuf, SolveCtl*, hipStream_t, bool, bool);
template void launch_mg_sweep_dots<float>(const MgLevel&, const float*, const float*, double*, const double*,
                                          const double*, RedBuf, SolveCtl*, hipStream_t, bool, bool);
void launch_fcg_norm(const MgLevel& L, double* rn, MeanRef mrn, const SolveCtl* ctl, RedBuf rb, hipStream_t s,
                     float* b0, float* x1f, double* x1d) {
  if (x1f && x1d) throw std::logic_error("fcg_norm: one first-sweep output");
  withCells(L, [&](auto nc, const Geo3& g) {
    if (x1f)
      k_fcg_norm<float, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, rn, mrn, ctl, rb, b0, x1f);
    else
      k_fcg_norm<double, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, rn, mrn, ctl, rb, b0, x1d);
  });
  NBL_LAUNCH_CHECK();
}
void launch_fcg_dots(const MgLevel& L, double* z, MeanRef mz, const double* rn, const double* r,
                     SolveCtl* ctl, RedBuf rb, hipStream_t s, bool countIt) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fcg_dots<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, z, mz, rn, r, ctl, rb, countIt ? 1 : 0);
  });
  NBL_LAUNCH_CHECK();
}
template <class T>
void launch_fcg_dots_x(const MgLevel& L, const T* x, MeanRef mx, double* z, const double* rn, const double* r,
                       SolveCtl* ctl, RedBuf rb, hipStream_t s, bool countIt) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fcg_dots_x<T, decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, x, mx, z, rn, r, ctl, rb, countIt ? 1 : 0);
  });
  NBL_LAUNCH_CHECK();
}
template void launch_fcg_dots_x<double>(const MgLevel&, const double*, MeanRef, double*, const double*,
                                        const double*, SolveCtl*, RedBuf, hipStream_t, bool);
template void launch_fcg_dots_x<float>(const MgLevel&, const float*, MeanRef, double*, const double*,
                                       const double*, SolveCtl*, RedBuf, hipStream_t, bool);
void launch_fluid_sum(const MgLevel& L, const double* x, RedBuf rb, hipStream_t s) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fluid_sum<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, x, rb);
  });
  NBL_LAUNCH_CHECK();
}
void launch_fcg_final(const MgLevel& L, bool singular, const double* x, const double* rhs, double* xo,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_fcg_final<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, singular ? 1 : 0, x, rhs, xo, ctl, rb);
  });
  NBL_LAUNCH_CHECK();
}
void launch_cg_z(const MgLevel& L, const double* r, double* z, bool first, double* p, const SolveCtl* ctl,
                 RedBuf rb, hipStream_t s) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_cg_z<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, r, z, p, first ? 1 : 0, ctl, rb);
  });
  NBL_LAUNCH_CHECK();
}
void launch_cg_update(const MgLevel& L, double* x, const double* p, double* r, const double* Ap,
                      const SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  withCells(L, [&](auto nc, const Geo3& g) {
    k_cg_update<decltype(nc)::value><<<g.grid, g.block, 0, s>>>(L, x, p, r, Ap, ctl, rb);
  });
  NBL_LAUNCH_CHECK();
}

}  // namespace nbl
