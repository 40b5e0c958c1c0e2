// mg.hip — gfx950 kernels of the pressure Poisson solve
// (reference core/src/flow/poisson.cpp and core/src/flow/multigrid.cpp).
//
// Memory model: one compact array per vector and level (i fastest) plus a
// 16-bit per-cell coupling code; the 7-point operator, its diagonal and the
// periodic wrap are all rebuilt from the code, so a sweep moves 3 vectors and
// 2 bytes per cell.  The reference's null-space projection ("removeMean",
// multigrid.cpp:81-99) is applied lazily: the kernel that produces a vector
// also folds its fluid sum, and every consumer subtracts the resulting mean
// when it reads the vector.  x - mean is the same rounded value whether it
// is stored first or formed on the fly, so this is bit-identical per element
// to the reference while saving one pass and one launch per projection.
//
// Every kernel of a flexible-CG iteration first checks ctl->done, so the host
// can enqueue iterations before it knows whether the solve has converged.
#include "dev.hpp"

namespace nbl {
namespace {

__device__ __forceinline__ long long nbIndex(const MgLevel& L, long long c, int i, int j, int k, int dir) {
  const long long sx = 1, sy = L.nx, sz = (long long)L.nx * L.ny;
  switch (dir) {
    case 0: return i > 0 ? c - sx : c + (L.nx - 1) * sx;
    case 1: return i < L.nx - 1 ? c + sx : c - (L.nx - 1) * sx;
    case 2: return j > 0 ? c - sy : c + (L.ny - 1) * sy;
    case 3: return j < L.ny - 1 ? c + sy : c - (L.ny - 1) * sy;
    case 4: return k > 0 ? c - sz : c + (L.nz - 1) * sz;
    default: return k < L.nz - 1 ? c + sz : c - (L.nz - 1) * sz;
  }
}

__device__ __forceinline__ void decode(const MgLevel& L, long long c, int& i, int& j, int& k) {
  i = int(c % L.nx);
  const long long q = c / L.nx;
  j = int(q % L.ny);
  k = int(q / L.ny);
}

// Value of x at a fluid cell as the reference holds it at that point of the
// cycle: x - pending mean (+ the parent's coarse correction after prolongation).
template <class T>
struct XView {
  const T* __restrict__ x;
  T m;
  bool hasM;
  const T* __restrict__ ec;  // non-null: add the parent's correction ec[parent] - em
  int ncx, ncy, fx, fy, fz;
  T em;
  bool hasEm;
  __device__ __forceinline__ T at(long long c, int i, int j, int k) const {
    T v = x[c];
    if (hasM) v = v - m;
    if (ec) {
      T e = ec[(i / fx) + (long long)ncx * ((j / fy) + (long long)ncy * (k / fz))];
      if (hasEm) e = e - em;
      v = v + e;
    }
    return v;
  }
};

__device__ __forceinline__ void stepTo(int dir, int& i, int& j, int& k, const MgLevel& L) {
  switch (dir) {
    case 0: i = i > 0 ? i - 1 : L.nx - 1; break;
    case 1: i = i < L.nx - 1 ? i + 1 : 0; break;
    case 2: j = j > 0 ? j - 1 : L.ny - 1; break;
    case 3: j = j < L.ny - 1 ? j + 1 : 0; break;
    case 4: k = k > 0 ? k - 1 : L.nz - 1; break;
    default: k = k < L.nz - 1 ? k + 1 : 0; break;
  }
}

// (A x)_c for a fluid cell, and the diagonal sum dg (poisson.cpp:102-134).
template <class T, bool ZERO>
__device__ __forceinline__ T applyRow(const MgLevel& L, uint16_t code, long long c, int i, int j, int k,
                                      const XView<T>& X, T& dg) {
  const T h2[3] = {T(L.h2[0]), T(L.h2[1]), T(L.h2[2])};
  T d = T(0), off = T(0);
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    const int kd = kindOf(code, dir);
    const int a = dir >> 1;
    if (kd == 1) {
      d += h2[a];
      if (!ZERO) {
        int ii = i, jj = j, kk = k;
        stepTo(dir, ii, jj, kk, L);
        off += h2[a] * X.at(nbIndex(L, c, i, j, k, dir), ii, jj, kk);
      } else {
        off += h2[a] * T(0);
      }
    } else if (kd == 2) {
      d += T(2) * h2[a];
    }
  }
  dg = d;
  const T xc = ZERO ? T(0) : X.at(c, i, j, k);
  return d * xc - off;
}

// Coarsest level: Jacobi-PCG to 1e-6 in at most 500 iterations, as
// CgPoisson(1e-6, 500).solve (multigrid.cpp:178-183, poisson.cpp:136-240),
// run by one workgroup with block-wide reductions.
constexpr int kCoarseThreads = 1024;

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
__global__ void __launch_bounds__(kCoarseThreads) k_coarse_cg(MgLevel L, int singular, const T* __restrict__ b,
                                                              const double* mb, T* __restrict__ x,
                                                              T* __restrict__ scr, const SolveCtl* ctl) {
  if (ctl && ctl->done) return;
  __shared__ double sh[2 * 16];
  __shared__ double bc[2];
  const long long N = L.n;
  T* r = scr;
  T* z = scr + N;
  T* p = scr + 2 * N;
  T* Ap = scr + 3 * N;
  const T bm = mb ? T(*mb) : T(0);
  const int tid = threadIdx.x, nt = blockDim.x;
  // rhs (kept in r), projected for singular systems
  double s1[1] = {0.0};
  for (long long c = tid; c < N; c += nt) {
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
    for (long long c = tid; c < N; c += nt) r[c] = (L.code[c] & kFluidBit) ? r[c] - m : T(0);
  }
  __syncthreads();
  const XView<T> P{p, T(0), false, nullptr, 1, 1, 1, 1, 1, T(0), false};
  double s2[2] = {0.0, 0.0};
  for (long long c = tid; c < N; c += nt) {
    const uint16_t code = L.code[c];
    T dg = T(0);
    if (code & kFluidBit) {
      int i, j, k;
      decode(L, c, i, j, k);
      const XView<T> Z{r, T(0), false, nullptr, 1, 1, 1, 1, 1, T(0), false};
      applyRow<T, true>(L, code, c, i, j, k, Z, dg);
    }
    const T diag = (code & kFluidBit) ? (dg > T(0) ? dg : T(1)) : T(1);
    const T zv = r[c] / diag;
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
    for (long long c = tid; c < N; c += nt) {
      const uint16_t code = L.code[c];
      T v;
      if (code & kFluidBit) {
        int i, j, k;
        decode(L, c, i, j, k);
        T dg;
        v = applyRow<T, false>(L, code, c, i, j, k, P, dg);
      } else {
        v = p[c];
      }
      Ap[c] = v;
      s[0] += double(p[c]) * double(v);
    }
    blockAll<1>(s, sh, bc);
    const double pAp = s[0];
    if (fabs(pAp) < 1e-300) break;
    const double alpha = rz / pAp;
    double q2[2] = {0.0, 0.0};
    for (long long c = tid; c < N; c += nt) {
      const uint16_t code = L.code[c];
      x[c] = x[c] + T(alpha) * p[c];
      const T rv = r[c] - T(alpha) * Ap[c];
      r[c] = rv;
      T dg = T(0);
      if (code & kFluidBit) {
        int i, j, k;
        decode(L, c, i, j, k);
        const XView<T> Z{r, T(0), false, nullptr, 1, 1, 1, 1, 1, T(0), false};
        applyRow<T, true>(L, code, c, i, j, k, Z, dg);
      }
      const T diag = (code & kFluidBit) ? (dg > T(0) ? dg : T(1)) : T(1);
      const T zv = rv / diag;
      z[c] = zv;
      q2[0] += double(rv) * double(rv);
      q2[1] += double(rv) * double(zv);
    }
    blockAll<2>(q2, sh, bc);
    if (sqrt(q2[0]) <= 1e-6 * bn) break;
    const double beta = q2[1] / rz;
    rz = q2[1];
    for (long long c = tid; c < N; c += nt) p[c] = z[c] + T(beta) * p[c];
    __syncthreads();
  }
  if (singular && L.nf > 0) {
    double sx[1] = {0.0};
    for (long long c = tid; c < N; c += nt)
      if (L.code[c] & kFluidBit) sx[0] += double(x[c]);
    blockAll<1>(sx, sh, bc);
    const T m = T(sx[0] / double(L.nf));
    for (long long c = tid; c < N; c += nt)
      if (L.code[c] & kFluidBit) x[c] = x[c] - m;
  }
}

// Level-0 preconditioner output: z = x - pending mean (fluid), 0 elsewhere;
// folds the fluid sum of z for Minv's own projection (multigrid.cpp:283-289).
template <class T>
__global__ void __launch_bounds__(kBlock) k_finish(MgLevel L, const T* __restrict__ x, const double* mx,
                                                   double* __restrict__ z, double* meanOut, RedBuf rb,
                                                   const SolveCtl* ctl) {
  if (ctl && ctl->done) return;
  const T m = mx ? T(*mx) : T(0);
  double acc[1] = {0.0};
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < L.n; c += (long long)gridDim.x * blockDim.x) {
    const bool fl = L.code[c] & kFluidBit;
    T v = x[c];
    if (mx && fl) v = v - m;
    const double zv = fl ? double(v) : 0.0;
    z[c] = zv;
    if (fl) acc[0] += zv;
  }
  if (!meanOut) return;
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) *meanOut = tot[0] / double(L.nf);
}

template <class T>
__global__ void k_to_level0(long long n, const double* __restrict__ r, T* __restrict__ b, const SolveCtl* ctl) {
  if (ctl && ctl->done) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) b[c] = T(r[c]);
}

// ---------------------------------------------------------------------------
// Plain operator pieces (kernel-level parity hooks)
// ---------------------------------------------------------------------------
__global__ void k_apply(MgLevel L, const double* __restrict__ x, double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.n) return;
  const uint16_t code = L.code[c];
  if (!(code & kFluidBit)) {
    out[c] = x[c];
    return;
  }
  int i, j, k;
  decode(L, c, i, j, k);
  const XView<double> X{x, 0.0, false, nullptr, 1, 1, 1, 1, 1, 0.0, false};
  double dg;
  out[c] = applyRow<double, false>(L, code, c, i, j, k, X, dg);
}

__global__ void k_diag(MgLevel L, double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.n) return;
  const uint16_t code = L.code[c];
  if (!(code & kFluidBit)) {
    out[c] = 1.0;
    return;
  }
  int i, j, k;
  decode(L, c, i, j, k);
  const XView<double> X{nullptr, 0.0, false, nullptr, 1, 1, 1, 1, 1, 0.0, false};
  double dg;
  applyRow<double, true>(L, code, c, i, j, k, X, dg);
  out[c] = dg > 0.0 ? dg : 1.0;
}

__global__ void k_prolong(MgLevel L, MgLevel Lc, const double* __restrict__ ec, double* __restrict__ xf) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.n) return;
  if (!(L.code[c] & kFluidBit)) return;
  int i, j, k;
  decode(L, c, i, j, k);
  xf[c] = xf[c] + ec[(i / Lc.fx) + (long long)Lc.nx * ((j / Lc.fy) + (long long)Lc.ny * (k / Lc.fz))];
}

// ---------------------------------------------------------------------------
// Flexible CG (MgPoisson::solve, multigrid.cpp:206-355) and Jacobi CG
// (CgPoisson::solve, poisson.cpp:136-252).  Scalars live in SolveCtl.
// ---------------------------------------------------------------------------
enum : int { S_MEAN_RHS = 0, S_MEAN_RN = 1, S_MEAN_X = 2 };

#define GRID_LOOP(c, n) \
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < (n); c += (long long)gridDim.x * blockDim.x)

__global__ void __launch_bounds__(kBlock) k_fcg_init(MgLevel L, int singular, int maskSolid, const double* __restrict__ b,
                                                     double* __restrict__ rhs, double* __restrict__ x,
                                                     SolveCtl* ctl, RedBuf rb) {
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    const bool fl = L.code[c] & kFluidBit;
    const double v = (fl || !maskSolid) ? b[c] : 0.0;
    rhs[c] = v;
    x[c] = 0.0;
    if (fl) acc[0] += v;
  }
  if (!singular) return;
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) ctl->slot[S_MEAN_RHS] = tot[0] / double(L.nf);
}

__global__ void __launch_bounds__(kBlock) k_fcg_init2(MgLevel L, int singular, int earlyExit, double* __restrict__ rhs,
                                                      double* __restrict__ r, SolveCtl* ctl, RedBuf rb) {
  const double m = singular ? ctl->slot[S_MEAN_RHS] : 0.0;
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    double v = rhs[c];
    if (singular) {
      v = (L.code[c] & kFluidBit) ? v - m : 0.0;
      rhs[c] = v;
    }
    r[c] = v;
    acc[0] += v * v;
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) {
    const double bn = sqrt(tot[0]) + 1e-300;
    ctl->bnorm = bn;
    ctl->it = 0;
    ctl->done = 0;
    ctl->converged = 0;
    ctl->broke = 0;
    ctl->vcycles = 0;
    if (earlyExit && sqrt(tot[0]) <= ctl->tol * bn) {  // multigrid.cpp:292-298
      ctl->done = 1;
      ctl->converged = 1;
      ctl->resid = sqrt(tot[0]) / bn;
      ctl->slot[63] = 1.0;  // marks "returned before the loop"
    } else {
      ctl->slot[63] = 0.0;
    }
  }
}

// z <- z - pending mean (fluid); p = z; rz = r.z
__global__ void __launch_bounds__(kBlock) k_fcg_start(MgLevel L, double* __restrict__ z, const double* mz,
                                                      const double* __restrict__ r, double* __restrict__ p,
                                                      SolveCtl* ctl, RedBuf rb) {
  if (ctl->done) return;
  const double m = mz ? *mz : 0.0;
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    double v = z[c];
    if (mz && (L.code[c] & kFluidBit)) v = v - m;
    z[c] = v;
    p[c] = v;
    acc[0] += r[c] * v;
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) ctl->rz = tot[0];
}

__global__ void __launch_bounds__(kBlock) k_fcg_apply(MgLevel L, const double* __restrict__ p,
                                                      double* __restrict__ Ap, SolveCtl* ctl, RedBuf rb) {
  if (ctl->done) return;
  const XView<double> X{p, 0.0, false, nullptr, 1, 1, 1, 1, 1, 0.0, false};
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    const uint16_t code = L.code[c];
    double v;
    if (code & kFluidBit) {
      int i, j, k;
      decode(L, c, i, j, k);
      double dg;
      v = applyRow<double, false>(L, code, c, i, j, k, X, dg);
    } else {
      v = p[c];
    }
    Ap[c] = v;
    acc[0] += p[c] * v;
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) {
    ctl->pAp = tot[0];
    if (fabs(tot[0]) < 1e-300) {
      ctl->done = 1;
      ctl->broke = 1;
    } else {
      ctl->alpha = ctl->rz / tot[0];
    }
  }
}

__device__ __forceinline__ void convergeCheck(SolveCtl* ctl, double rr) {
  if (sqrt(rr) <= ctl->tol * ctl->bnorm) {
    ctl->it += 1;
    ctl->converged = 1;
    ctl->done = 1;
  }
}

// x += alpha p; rn = r - alpha Ap; fold sum(rn) (singular) or |rn|^2.
__global__ void __launch_bounds__(kBlock) k_fcg_update(MgLevel L, int singular, double* __restrict__ x,
                                                       const double* __restrict__ p, const double* __restrict__ r,
                                                       const double* __restrict__ Ap, double* __restrict__ rn,
                                                       SolveCtl* ctl, RedBuf rb) {
  if (ctl->done) return;
  const double a = ctl->alpha;
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    x[c] = x[c] + a * p[c];
    const double v = r[c] - a * Ap[c];
    rn[c] = v;
    if (singular) {
      if (L.code[c] & kFluidBit) acc[0] += v;
    } else {
      acc[0] += v * v;
    }
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) {
    if (singular) ctl->slot[S_MEAN_RN] = tot[0] / double(L.nf);
    else convergeCheck(ctl, tot[0]);
  }
}

__global__ void __launch_bounds__(kBlock) k_fcg_norm(MgLevel L, double* __restrict__ rn, SolveCtl* ctl, RedBuf rb) {
  if (ctl->done) return;
  const double m = ctl->slot[S_MEAN_RN];
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    const double v = (L.code[c] & kFluidBit) ? rn[c] - m : 0.0;
    rn[c] = v;
    acc[0] += v * v;
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) convergeCheck(ctl, tot[0]);
}

// z <- z - pending mean; a1 = z.rn, a2 = z.r; beta = (a1 - a2)/rz (flexible,
// Polak-Ribiere, multigrid.cpp:328-339).
__global__ void __launch_bounds__(kBlock) k_fcg_dots(MgLevel L, double* __restrict__ z, const double* mz,
                                                     const double* __restrict__ rn, const double* __restrict__ r,
                                                     SolveCtl* ctl, RedBuf rb, int maxIt) {
  if (ctl->done) return;
  const double m = mz ? *mz : 0.0;
  double acc[2] = {0.0, 0.0};
  GRID_LOOP(c, L.n) {
    double v = z[c];
    if (mz && (L.code[c] & kFluidBit)) v = v - m;
    z[c] = v;
    acc[0] += v * rn[c];
    acc[1] += v * r[c];
  }
  double tot[2];
  if (gridFold<2>(acc, rb, tot) && threadIdx.x == 0) {
    ctl->beta = (tot[0] - tot[1]) / ctl->rz;
    ctl->rz = tot[0];
    ctl->it += 1;
    if (ctl->it >= maxIt) ctl->done = 1;
  }
}

__global__ void __launch_bounds__(kBlock) k_fcg_pupdate(long long n, double* __restrict__ p, const double* __restrict__ z,
                                                        double* __restrict__ r, const double* __restrict__ rn,
                                                        SolveCtl* ctl) {
  if (ctl->done) return;
  const double be = ctl->beta;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  p[c] = z[c] + be * p[c];
  if (r) r[c] = rn[c];
}

__global__ void __launch_bounds__(kBlock) k_fluid_sum(MgLevel L, const double* __restrict__ x, SolveCtl* ctl,
                                                      RedBuf rb) {
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) if (L.code[c] & kFluidBit) acc[0] += x[c];
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) ctl->slot[S_MEAN_X] = tot[0] / double(L.nf);
}

// Final projection of x and the true relative residual (multigrid.cpp:342-354).
__global__ void __launch_bounds__(kBlock) k_fcg_final(MgLevel L, int singular, const double* __restrict__ x,
                                                      const double* __restrict__ rhs, double* __restrict__ xo,
                                                      SolveCtl* ctl, RedBuf rb, double convTol) {
  const double m = singular ? ctl->slot[S_MEAN_X] : 0.0;
  const XView<double> X{x, m, singular != 0, nullptr, 1, 1, 1, 1, 1, 0.0, false};
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    const uint16_t code = L.code[c];
    double Ax, xv;
    if (code & kFluidBit) {
      int i, j, k;
      decode(L, c, i, j, k);
      double dg;
      Ax = applyRow<double, false>(L, code, c, i, j, k, X, dg);
      xv = singular ? x[c] - m : x[c];
    } else {
      xv = singular ? 0.0 : x[c];
      Ax = xv;
    }
    xo[c] = xv;
    const double e = rhs[c] - Ax;
    acc[0] += e * e;
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) {
    if (ctl->slot[63] != 0.0) return;  // returned before the loop: keep that report
    ctl->resid = sqrt(tot[0]) / ctl->bnorm;
    ctl->converged = (ctl->converged || ctl->resid <= convTol) ? 1 : 0;
  }
}

// Jacobi-CG pieces: z = r/diag and its dot with r.
__global__ void __launch_bounds__(kBlock) k_cg_z(MgLevel L, const double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, SolveCtl* ctl, RedBuf rb, int first,
                                                 int maxIt) {
  if (ctl->done) return;
  const XView<double> X{nullptr, 0.0, false, nullptr, 1, 1, 1, 1, 1, 0.0, false};
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    const uint16_t code = L.code[c];
    double diag = 1.0;
    if (code & kFluidBit) {
      int i, j, k;
      decode(L, c, i, j, k);
      double dg;
      applyRow<double, true>(L, code, c, i, j, k, X, dg);
      diag = dg > 0.0 ? dg : 1.0;
    }
    const double zv = r[c] / diag;
    z[c] = zv;
    if (first) p[c] = zv;
    acc[0] += r[c] * zv;
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) {
    if (first) {
      ctl->rz = tot[0];
    } else {
      ctl->beta = tot[0] / ctl->rz;
      ctl->rz = tot[0];
      ctl->it += 1;
      if (ctl->it >= maxIt) ctl->done = 1;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_cg_update(MgLevel L, double* __restrict__ x, const double* __restrict__ p,
                                                      double* __restrict__ r, const double* __restrict__ Ap,
                                                      SolveCtl* ctl, RedBuf rb) {
  if (ctl->done) return;
  const double a = ctl->alpha;
  double acc[1] = {0.0};
  GRID_LOOP(c, L.n) {
    x[c] = x[c] + a * p[c];
    const double v = r[c] - a * Ap[c];
    r[c] = v;
    acc[0] += v * v;
  }
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) convergeCheck(ctl, tot[0]);
}

int redGrid(long long n) { return gridFor(n, kBlock, kMaxRedBlocks); }

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// Smoother: x <- x + 0.8 (b - A x)/diag on fluid cells, 0 elsewhere
// (multigrid.cpp:102-114), with the fused prolongation of the parent
// correction (multigrid.cpp:158-173) when `ec` is given, and the fluid sum of
// the result for the projection that follows a sweep on singular systems.
// Pending means (`mb`, `mx`) are device slots, read here.
template <class T, bool ZERO>
__global__ void __launch_bounds__(kBlock) k_sweep_w(MgLevel L, const T* __restrict__ b, const double* mb,
                                                    const T* __restrict__ x, const double* mx, const T* ec,
                                                    const double* mec, int ncx, int ncy, int fx, int fy, int fz,
                                                    T* __restrict__ xo, double* meanOut, RedBuf rb,
                                                    const SolveCtl* ctl) {
  if (ctl && ctl->done) return;
  XView<T> X{x, mx ? T(*mx) : T(0), mx != nullptr, ec, ncx, ncy, fx, fy, fz, mec ? T(*mec) : T(0), mec != nullptr};
  const T bm = mb ? T(*mb) : T(0);
  double acc[1] = {0.0};
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < L.n; c += (long long)gridDim.x * blockDim.x) {
    const uint16_t code = L.code[c];
    if (!(code & kFluidBit)) {
      xo[c] = T(0);
      continue;
    }
    int i, j, k;
    decode(L, c, i, j, k);
    T dg;
    const T Ax = applyRow<T, ZERO>(L, code, c, i, j, k, X, dg);
    const T xc = ZERO ? T(0) : X.at(c, i, j, k);
    const T diag = dg > T(0) ? dg : T(1);
    T bv = b[c];
    if (mb) bv = bv - bm;
    const T xn = xc + T(0.8) * (bv - Ax) / diag;
    xo[c] = xn;
    acc[0] += double(xn);
  }
  if (!meanOut) return;
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) *meanOut = tot[0] / double(L.nf);
}

// Residual r = b - A x on fine fluid cells restricted to the coarse grid as
// the mean over fluid children, children summed in the reference's k,j,i
// order (multigrid.cpp:116-154).  ZEROX: x == 0 (restriction of b itself).
template <class T, bool ZEROX>
__global__ void __launch_bounds__(kBlock) k_resid_restrict_w(MgLevel L, MgLevel Lc, const T* __restrict__ b,
                                                             const double* mb, const T* __restrict__ x,
                                                             const double* mx, T* __restrict__ rc, double* meanOut,
                                                             RedBuf rb, const SolveCtl* ctl) {
  if (ctl && ctl->done) return;
  XView<T> X{x, mx ? T(*mx) : T(0), mx != nullptr, nullptr, 1, 1, 1, 1, 1, T(0), false};
  const T bm = mb ? T(*mb) : T(0);
  double acc[1] = {0.0};
  for (long long C = (long long)blockIdx.x * blockDim.x + threadIdx.x; C < Lc.n; C += (long long)gridDim.x * blockDim.x) {
    if (!(Lc.code[C] & kFluidBit)) {
      rc[C] = T(0);
      continue;
    }
    int I, J, K;
    decode(Lc, C, I, J, K);
    T s = T(0);
    int cnt = 0;
    const int k1 = min((K + 1) * Lc.fz, L.nz), j1 = min((J + 1) * Lc.fy, L.ny), i1 = min((I + 1) * Lc.fx, L.nx);
    for (int k = K * Lc.fz; k < k1; ++k)
      for (int j = J * Lc.fy; j < j1; ++j)
        for (int i = I * Lc.fx; i < i1; ++i) {
          const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
          const uint16_t code = L.code[c];
          if (!(code & kFluidBit)) continue;
          T dg;
          const T Ax = applyRow<T, ZEROX>(L, code, c, i, j, k, X, dg);
          T bv = b[c];
          if (mb) bv = bv - bm;
          s += bv - Ax;
          ++cnt;
        }
    const T r = s / T(cnt);
    rc[C] = r;
    acc[0] += double(r);
  }
  if (!meanOut) return;
  double tot[1];
  if (gridFold<1>(acc, rb, tot) && threadIdx.x == 0) *meanOut = tot[0] / double(Lc.nf);
}

template <class T>
void launch_mg_sweep(const MgLevel& L, const T* b, const double* mb, const T* x, const double* mx,
                     bool zeroGuess, const T* ec, const double* mec, const MgLevel* Lc, T* xo,
                     double* meanOut, RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  const int g = redGrid(L.n);
  const int ncx = Lc ? Lc->nx : 1, ncy = Lc ? Lc->ny : 1;
  const int fx = Lc ? Lc->fx : 1, fy = Lc ? Lc->fy : 1, fz = Lc ? Lc->fz : 1;
  if (zeroGuess)
    k_sweep_w<T, true><<<g, kBlock, 0, s>>>(L, b, mb, x, mx, nullptr, nullptr, 1, 1, 1, 1, 1, xo, meanOut, rb,
                                            ctl);
  else
    k_sweep_w<T, false><<<g, kBlock, 0, s>>>(L, b, mb, x, mx, ec, mec, ncx, ncy, fx, fy, fz, xo, meanOut, rb,
                                             ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_resid_restrict(const MgLevel& L, const MgLevel& Lc, const T* b, const double* mb,
                              const T* x, const double* mx, T* rc, double* meanOut, RedBuf rb,
                              const SolveCtl* ctl, hipStream_t s) {
  k_resid_restrict_w<T, false><<<redGrid(Lc.n), kBlock, 0, s>>>(L, Lc, b, mb, x, mx, rc, meanOut, rb, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_coarse(const MgLevel& L, bool singular, const T* b, const double* mb, T* x,
                      T* scratch, const SolveCtl* ctl, hipStream_t s) {
  k_coarse_cg<T><<<1, kCoarseThreads, 0, s>>>(L, singular ? 1 : 0, b, mb, x, scratch, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_mg_finish(const MgLevel& L, const T* x, const double* mx, double* z, double* meanOut,
                      RedBuf rb, const SolveCtl* ctl, hipStream_t s) {
  k_finish<T><<<redGrid(L.n), kBlock, 0, s>>>(L, x, mx, z, meanOut, rb, ctl);
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_to_level0(const MgLevel& L, const double* r, T* b, const SolveCtl* ctl, hipStream_t s) {
  k_to_level0<T><<<gridFor(L.n), kBlock, 0, s>>>(L.n, r, b, ctl);
  NBL_LAUNCH_CHECK();
}

#define NBL_INST(T)                                                                                        \
  template void launch_mg_sweep<T>(const MgLevel&, const T*, const double*, const T*, const double*, bool,  \
                                   const T*, const double*, const MgLevel*, T*, double*, RedBuf,            \
                                   const SolveCtl*, hipStream_t);                                          \
  template void launch_mg_resid_restrict<T>(const MgLevel&, const MgLevel&, const T*, const double*,       \
                                            const T*, const double*, T*, double*, RedBuf, const SolveCtl*, \
                                            hipStream_t);                                                  \
  template void launch_mg_coarse<T>(const MgLevel&, bool, const T*, const double*, T*, T*,                 \
                                    const SolveCtl*, hipStream_t);                                         \
  template void launch_mg_finish<T>(const MgLevel&, const T*, const double*, double*, double*, RedBuf,     \
                                    const SolveCtl*, hipStream_t);                                         \
  template void launch_to_level0<T>(const MgLevel&, const double*, T*, const SolveCtl*, hipStream_t);
NBL_INST(double)
NBL_INST(float)
#undef NBL_INST

void launch_apply(const MgLevel& L, const double* x, double* out, hipStream_t s) {
  k_apply<<<gridFor(L.n), kBlock, 0, s>>>(L, x, out);
  NBL_LAUNCH_CHECK();
}
void launch_diag(const MgLevel& L, double* out, hipStream_t s) {
  k_diag<<<gridFor(L.n), kBlock, 0, s>>>(L, out);
  NBL_LAUNCH_CHECK();
}
void launch_restrict_plain(const MgLevel& L, const MgLevel& Lc, const double* rf, double* rc, hipStream_t s) {
  // rf - A*0 == rf exactly, so the fused kernel with a zero x restricts rf itself
  k_resid_restrict_w<double, true><<<redGrid(Lc.n), kBlock, 0, s>>>(L, Lc, rf, nullptr, nullptr, nullptr, rc, nullptr,
                                                                RedBuf{nullptr, nullptr}, nullptr);
  NBL_LAUNCH_CHECK();
}
void launch_prolong_plain(const MgLevel& L, const MgLevel& Lc, const double* ec, double* xf, hipStream_t s) {
  k_prolong<<<gridFor(L.n), kBlock, 0, s>>>(L, Lc, ec, xf);
  NBL_LAUNCH_CHECK();
}

void launch_fcg_init(const MgLevel& L, bool singular, bool maskSolid, const double* b, double* rhs,
                     double* x, SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  k_fcg_init<<<redGrid(L.n), kBlock, 0, s>>>(L, singular ? 1 : 0, maskSolid ? 1 : 0, b, rhs, x, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_init2(const MgLevel& L, bool singular, bool earlyExit, double* rhs, double* r,
                      SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  k_fcg_init2<<<redGrid(L.n), kBlock, 0, s>>>(L, singular ? 1 : 0, earlyExit ? 1 : 0, rhs, r, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_start(const MgLevel& L, const double* z, const double* mz, const double* r,
                      double* zm, double* p, SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  (void)z;
  k_fcg_start<<<redGrid(L.n), kBlock, 0, s>>>(L, zm, mz, r, p, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_apply(const MgLevel& L, const double* p, double* Ap, SolveCtl* ctl, RedBuf rb,
                      hipStream_t s) {
  k_fcg_apply<<<redGrid(L.n), kBlock, 0, s>>>(L, p, Ap, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_update(const MgLevel& L, bool singular, double* x, const double* p,
                       const double* r, const double* Ap, double* rn, SolveCtl* ctl, RedBuf rb,
                       hipStream_t s) {
  k_fcg_update<<<redGrid(L.n), kBlock, 0, s>>>(L, singular ? 1 : 0, x, p, r, Ap, rn, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_norm(const MgLevel& L, double* rn, SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  k_fcg_norm<<<redGrid(L.n), kBlock, 0, s>>>(L, rn, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_dots(const MgLevel& L, double* z, const double* mz, const double* rn,
                     const double* r, SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  k_fcg_dots<<<redGrid(L.n), kBlock, 0, s>>>(L, z, mz, rn, r, ctl, rb, 1 << 30);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_pupdate(const MgLevel& L, double* p, const double* z, double* r,
                        const double* rn, SolveCtl* ctl, hipStream_t s) {
  k_fcg_pupdate<<<gridFor(L.n), kBlock, 0, s>>>(L.n, p, z, r, rn, ctl);
  NBL_LAUNCH_CHECK();
}
void launch_fcg_final(const MgLevel& L, bool singular, const double* x, const double* rhs,
                      double* xo, SolveCtl* ctl, RedBuf rb, double convTol, hipStream_t s) {
  if (singular) {
    k_fluid_sum<<<redGrid(L.n), kBlock, 0, s>>>(L, x, ctl, rb);
    NBL_LAUNCH_CHECK();
  }
  k_fcg_final<<<redGrid(L.n), kBlock, 0, s>>>(L, singular ? 1 : 0, x, rhs, xo, ctl, rb, convTol);
  NBL_LAUNCH_CHECK();
}
void launch_cg_z(const MgLevel& L, const double* r, double* z, SolveCtl* ctl, RedBuf rb,
                 bool first, double* p, hipStream_t s) {
  k_cg_z<<<redGrid(L.n), kBlock, 0, s>>>(L, r, z, p, ctl, rb, first ? 1 : 0, 1 << 30);
  NBL_LAUNCH_CHECK();
}
void launch_cg_update(const MgLevel& L, double* x, const double* p, double* r, const double* Ap,
                      SolveCtl* ctl, RedBuf rb, hipStream_t s) {
  k_cg_update<<<redGrid(L.n), kBlock, 0, s>>>(L, x, p, r, Ap, ctl, rb);
  NBL_LAUNCH_CHECK();
}
void launch_cg_pupdate(const MgLevel& L, double* p, const double* z, SolveCtl* ctl, hipStream_t s) {
  k_fcg_pupdate<<<gridFor(L.n), kBlock, 0, s>>>(L.n, p, z, nullptr, nullptr, ctl);
  NBL_LAUNCH_CHECK();
}

}  // namespace nbl
