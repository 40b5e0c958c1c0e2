// vcycle.hip — the V(2,2) cycle of MgPoisson (multigrid.cpp:175-194) in two
// kernels per level.
//
// A grid-wide barrier inside one launch costs more on gfx950 than a kernel
// boundary does (measured: 2.6 us per back-to-back launch; 1-25 us per grid
// barrier over 32-512 workgroups, and more once the barrier must publish
// stores across the eight XCD L2s), so the cycle's launch count is cut by
// temporal blocking instead: each workgroup owns a tile of TX x TY x TZ cells,
// stages its level vectors over the tile plus two halo layers in LDS, and
// recomputes the halo sweeps itself.
//
//   down kernel (level l): the first pre-sweep from zero over tile + 2 halo,
//     the second pre-sweep over tile + 1 halo, the residual on the tile and
//     its restriction to the coarse cells the tile covers (multigrid.cpp:
//     102-154).  On level 0 the first sweep's right-hand side is formed on the
//     fly from the flexible CG's update r_new = r - alpha A p
//     (multigrid.cpp:314-317), whose x and r_new the kernel also writes.
//   up kernel (level l): the coarse correction prolonged over tile + 2 halo
//     (multigrid.cpp:158-173), the first post-sweep over tile + 1 halo, the
//     second on the tile; on level 0 the second writes the preconditioner
//     output z and the flexible CG's dots z.r_new, z.r.
//
// Every cell value is evaluated with the expressions, in the order, of the
// per-sweep kernels (k_sweep, k_resid_restrict, k_fcg_update_first,
// k_sweep_dots): on systems without null-space projections the cycle is
// bit-identical to them per element; only the partial-sum tree of the
// reductions differs.  On singular systems the reference removes the fluid
// mean after every sweep (multigrid.cpp:110-112); the operator annihilates
// constants on fluid cells (every coupling of a fluid row is to a fluid
// neighbour and no outlet exists), so those projections change neither the
// residual nor, after the final projection of z, the preconditioner's output
// in exact arithmetic.  Here they are folded into the projections that do
// change values: the restricted right-hand side of every level and r_new
// (multigrid.cpp:150-152, 318), and z (multigrid.cpp:288).
#include <stdexcept>
#include <type_traits>

#include "mgkern.hpp"

namespace nbl {
namespace {

__device__ __forceinline__ int wrapc(int v, int n) {
  while (v < 0) v += n;
  while (v >= n) v -= n;
  return v;
}

// Tile geometry: TX x TY x TZ cells; HZ = 2 on 3-D levels (two halo layers
// along z), 0 on 2-D levels (nz == 1, no z couplings); KT threads per block.
constexpr int KT = 512;
template <int TX, int TY, int TZ, int HZ>
struct Tile {
  static constexpr int RX2 = TX + 4, RY2 = TY + 4, RZ2 = TZ + 2 * HZ;  // tile + 2 halo
  static constexpr int RX1 = TX + 2, RY1 = TY + 2, RZ1 = TZ + HZ;      // tile + 1 halo
  static constexpr int N2 = RX2 * RY2 * RZ2, N1 = RX1 * RY1 * RZ1, NT = TX * TY * TZ;
  static constexpr int S2 = (N2 + KT - 1) / KT, S1 = (N1 + KT - 1) / KT;
  static constexpr int ST = (NT + KT - 1) / KT;
  static constexpr int P2 = RX2 * RY2, P1 = RX1 * RY1;
  // R2 slot -> R1 slot (-1 outside tile + 1)
  __device__ static int r1of(int ai, int bj, int ck) {
    const int a = ai - 1, b = bj - 1, c = ck - HZ / 2;
    return (a >= 0 && a < RX1 && b >= 0 && b < RY1 && c >= 0 && c < RZ1) ? a + RX1 * (b + RY1 * c) : -1;
  }
};

// The diagonal sum of a cell's row, accumulated in applyDecode's order
// (poisson.cpp:54-74): h2 for a fluid neighbour, 2 h2 for an outlet face.
template <class T>
__device__ __forceinline__ T diagSum(uint16_t code, const T (&h2)[3]) {
  T d = T(0);
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    const int kd = kindOf(code, dir);
    if (kd == 1) d += h2[dir >> 1];
    else if (kd == 2) d += T(2) * h2[dir >> 1];
  }
  return d;
}
// (A x) of a fluid row from its diagonal sum d (applyDecode's value: d x_c
// minus the fluid couplings summed in face order).
template <class T>
__device__ __forceinline__ T applyD(uint16_t code, T d, T xc, const T (&xn)[6], const T (&h2)[3]) {
  T off = T(0);
#pragma unroll
  for (int dir = 0; dir < 6; ++dir)
    if (kindOf(code, dir) == 1) off += h2[dir >> 1] * xn[dir];
  return d * xc - off;
}
// k_sweep's update at a fluid or solid cell
template <class T>
__device__ __forceinline__ T sweepD(uint16_t code, T d, T xc, const T (&xn)[6], T bv, const T (&h2)[3]) {
  if (!(code & kFluidBit)) return T(0);
  const T Ax = applyD<T>(code, d, xc, xn, h2);
  const T diag = d > T(0) ? d : T(1);
  return xc + T(0.8) * (bv - Ax) / diag;
}
// firstSweep's value: the sweep from a zero guess (couplings times zero)
template <class T>
__device__ __forceinline__ T firstD(uint16_t code, T d, T bv) {
  if (!(code & kFluidBit)) return T(0);
  const T Ax = d * T(0) - T(0);
  const T diag = d > T(0) ? d : T(1);
  return T(0) + T(0.8) * (bv - Ax) / diag;
}
// the six neighbours of LDS slot i in a box with row stride RX, plane stride P
template <class T, int RX, int P, bool ZDIM>
__device__ __forceinline__ void nbrs(const T* __restrict__ s, int i, T (&xn)[6]) {
  xn[0] = s[i - 1];
  xn[1] = s[i + 1];
  xn[2] = s[i - RX];
  xn[3] = s[i + RX];
  xn[4] = ZDIM ? s[i - P] : T(0);
  xn[5] = ZDIM ? s[i + P] : T(0);
}

// ---------------------------------------------------------------------------
// Down kernel
//
// Every global store waits until after the block's last barrier: a
// __syncthreads() is a workgroup-scope fence as well and would otherwise wait
// for the stores in flight.  All loads, including the tile's own x and p and
// the coarse cells' codes, are issued first, before the scalar folds.  Each
// cell's code is decoded into its diagonal sum once per kernel; the sweeps
// then accumulate only the couplings.
// ---------------------------------------------------------------------------
template <class T, int SRC, int TX, int TY, int TZ, int HZ, bool BF>
__global__ void __launch_bounds__(KT) k_down(DownArgs<T> a) {
  using G = Tile<TX, TY, TZ, HZ>;
  constexpr bool ZD = HZ > 0;
  constexpr bool UPD = SRC != DS_B;
  __shared__ T s1[G::N2];          // first sweep; the tile's residuals after phase B
  __shared__ T sb[G::N1];          // right-hand side, tile + 1
  __shared__ T sd[G::N1];          // diagonal sums, tile + 1
  __shared__ uint16_t sc[G::N1];   // coupling codes, tile + 1
  __shared__ T s2[G::N1];          // second sweep
  __shared__ unsigned char sfl[G::NT];
  const MgLevel& L = a.L;
  const MgLevel& Lc = a.Lc;
  const T h2[3] = {T(L.h2[0]), T(L.h2[1]), T(L.h2[2])};
  const int tid = threadIdx.x;
  const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY, k0 = blockIdx.z * TZ;
  const int cx = TX / Lc.fx, cy = TY / Lc.fy, cz = TZ / Lc.fz;

  // ---- loads: tile + 2 halo, the tile's own x and p, the coarse codes ----
  uint16_t cv[G::S2];
  double v0[G::S2], v1[G::S2], xo[G::S2], po[G::S2];
  int cidx[G::S2], r1[G::S2];
  bool own[G::S2];
#pragma unroll
  for (int q = 0; q < G::S2; ++q) {
    const int s = tid + q * KT;
    cidx[q] = 0;
    own[q] = false;
    r1[q] = -1;
    if (s < G::N2) {
      const int ai = s % G::RX2, t = s / G::RX2, bj = t % G::RY2, ck = t / G::RY2;
      const int ui = i0 - 2 + ai, uj = j0 - 2 + bj, uk = ZD ? k0 - HZ + ck : 0;
      const int gi = wrapc(ui, L.nx), gj = wrapc(uj, L.ny), gk = ZD ? wrapc(uk, L.nz) : 0;
      const int c = gi + L.nx * (gj + L.ny * gk);
      cidx[q] = c;
      r1[q] = G::r1of(ai, bj, ck);
      own[q] = ai >= 2 && ai < 2 + TX && bj >= 2 && bj < 2 + TY && ck >= HZ && ck < HZ + TZ && ui < L.nx &&
               uj < L.ny && uk < L.nz;
      cv[q] = L.code[c];
      if constexpr (SRC == DS_B) {
        v0[q] = double(a.b[c]);
      } else if constexpr (SRC == DS_UPD) {
        v0[q] = a.r[c];
        v1[q] = a.Ap[c];
        if (own[q]) {
          xo[q] = a.x[c];
          po[q] = a.p[c];
        }
      } else {
        v0[q] = a.rnRaw[c];
      }
    }
  }
  uint16_t ccv[G::ST];
  int cC[G::ST];
#pragma unroll
  for (int q = 0; q < G::ST; ++q) {
    const int u = tid + q * KT;
    cC[q] = -1;
    if (u < cx * cy * cz) {
      const int ci = u % cx, t = u / cx, cj = t % cy, ck = t / cy;
      const int I = blockIdx.x * cx + ci, J = blockIdx.y * cy + cj, K = blockIdx.z * cz + ck;
      if (I < Lc.nx && J < Lc.ny && K < Lc.nz) {
        cC[q] = I + Lc.nx * (J + Lc.ny * K);
        ccv[q] = Lc.code[cC[q]];
      }
    }
  }
  // ---- scalars: the convergence test, alpha, the pending mean ----
  // Block 0 of this kernel may stop the solve in the convergence test, so the
  // stop flag is read once per block (inside the folds) and not again after it.
  const SolveCtl* ctlM = a.ctl;
  if (a.conv.n) {
    if (inlineConv(a.conv, a.ctl)) return;
    ctlM = nullptr;
  }
  double alpha = 0.0, m = 0.0;
  if constexpr (SRC == DS_UPD) {
    if (a.alpha.n) {
      if (!inlineAlpha(a.alpha, a.ctl, alpha)) return;
    } else {
      if (ldA(&a.ctl->done)) return;
      alpha = ldA(&a.ctl->alpha);
    }
  } else {
    double mv[1];
    if (stopOrMeans<1>(ctlM, {SRC == DS_RN ? a.mrn : a.mb}, mv)) return;
    m = mv[0];
  }
  const bool hasM = SRC == DS_RN ? a.mrn.on() : a.mb.on();

  // ---- phase A: right-hand side and the first sweep from zero, tile + 2 ----
  double accN[1] = {0.0};
#pragma unroll
  for (int q = 0; q < G::S2; ++q) {
    const int s = tid + q * KT;
    if (s >= G::N2) continue;
    const uint16_t code = cv[q];
    const bool fl = code & kFluidBit;
    T bv;
    if constexpr (SRC == DS_B) {
      bv = T(v0[q]);
      if (hasM && fl) bv = bv - T(m);
    } else {
      const double rn = SRC == DS_UPD ? v0[q] - alpha * v1[q] : (fl ? v0[q] - m : 0.0);
      v0[q] = rn;  // r_new, stored after the last barrier
      bv = T(rn);
      if (own[q]) {
        accN[0] += rn * rn;
        if constexpr (SRC == DS_UPD) xo[q] = xo[q] + alpha * po[q];
      }
    }
    const T d = diagSum<T>(code, h2);
    s1[s] = firstD<T>(code, d, bv);
    if (r1[q] >= 0) {
      sb[r1[q]] = bv;
      sd[r1[q]] = d;
      sc[r1[q]] = code;
    }
  }
  __syncthreads();

  // ---- phase B: second pre-sweep over tile + 1 ----
  T xav[G::S1];
  int xac[G::S1];
#pragma unroll
  for (int q = 0; q < G::S1; ++q) {
    const int s = tid + q * KT;
    xac[q] = -1;
    if (s >= G::N1) continue;
    const int ai = s % G::RX1, t = s / G::RX1, bj = t % G::RY1, ck = t / G::RY1;
    const int i2 = (ai + 1) + G::RX2 * ((bj + 1) + G::RY2 * (ck + HZ / 2));
    T xn[6];
    nbrs<T, G::RX2, G::P2, ZD>(s1, i2, xn);
    const T out = sweepD<T>(sc[s], sd[s], s1[i2], xn, sb[s], h2);
    s2[s] = out;
    const int gi = i0 - 1 + ai, gj = j0 - 1 + bj, gk = ZD ? k0 - HZ / 2 + ck : 0;
    if (ai >= 1 && ai < 1 + TX && bj >= 1 && bj < 1 + TY && ck >= HZ / 2 && ck < HZ / 2 + TZ && gi < L.nx &&
        gj < L.ny && gk < L.nz) {
      xac[q] = gi + L.nx * (gj + L.ny * gk);
      xav[q] = out;
    }
  }
  __syncthreads();

  // ---- phase C: residual on the tile ----
  T* rs = s1;
#pragma unroll
  for (int q = 0; q < G::ST; ++q) {
    const int s = tid + q * KT;
    if (s >= G::NT) continue;
    const int ai = s % TX, t = s / TX, bj = t % TY, ck = t / TY;
    const int i1 = (ai + 1) + G::RX1 * ((bj + 1) + G::RY1 * (ck + HZ / 2));
    const uint16_t code = sc[i1];
    const bool fl = (code & kFluidBit) && i0 + ai < L.nx && j0 + bj < L.ny && k0 + ck < L.nz;
    sfl[s] = fl ? 1 : 0;
    T xn[6];
    nbrs<T, G::RX1, G::P1, ZD>(s2, i1, xn);
    rs[s] = fl ? sb[i1] - applyD<T>(code, sd[i1], s2[i1], xn, h2) : T(0);
  }
  __syncthreads();

  // ---- phase D: restriction (mean over fluid children, k, j, i order) ----
  double accM[1] = {0.0};
  T rcv[G::ST];
#pragma unroll
  for (int q = 0; q < G::ST; ++q) {
    rcv[q] = T(0);
    if (cC[q] < 0 || !(ccv[q] & kFluidBit)) continue;
    const int u = tid + q * KT;
    const int ci = u % cx, t = u / cx, cj = t % cy, ck = t / cy;
    T sum = T(0);
    int cnt = 0;
    for (int dk = 0; dk < Lc.fz; ++dk)
      for (int dj = 0; dj < Lc.fy; ++dj)
        for (int di = 0; di < Lc.fx; ++di) {
          const int tt = (ci * Lc.fx + di) + TX * ((cj * Lc.fy + dj) + TY * (ck * Lc.fz + dk));
          if (sfl[tt]) {
            sum += rs[tt];
            ++cnt;
          }
        }
    rcv[q] = sum / T(cnt);
    accM[0] += double(rcv[q]);
  }
  if (a.rowNorm) blockToPartials<1>(accN, a.rowNorm, 0, blk);
  if (a.rowMean) blockToPartials<1>(accM, a.rowMean, 0, blk);

  // ---- stores ----
  if constexpr (UPD) {
#pragma unroll
    for (int q = 0; q < G::S2; ++q) {
      if (!own[q]) continue;
      const int c = cidx[q];
      if constexpr (SRC == DS_UPD) a.x[c] = xo[q];
      a.rnOut[c] = v0[q];
      if (a.b0Out) a.b0Out[c] = T(v0[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < G::S1; ++q) {
    if (xac[q] < 0) continue;
    if constexpr (BF) a.xaB[xac[q]] = __float2bfloat16(float(xav[q]));
    else a.xa[xac[q]] = xav[q];
  }
#pragma unroll
  for (int q = 0; q < G::ST; ++q)
    if (cC[q] >= 0) a.bc[cC[q]] = rcv[q];
}

// ---------------------------------------------------------------------------
// Up kernel (stores after the last barrier, codes decoded once, as the down
// kernel)
// ---------------------------------------------------------------------------
template <class T, int DST, int TX, int TY, int TZ, int HZ, bool BF>
__global__ void __launch_bounds__(KT) k_up(UpArgs<T> a) {
  using G = Tile<TX, TY, TZ, HZ>;
  constexpr bool ZD = HZ > 0;
  __shared__ T s1[G::N2];  // x after the pre-sweeps + the prolonged correction
  __shared__ T sb[G::N1];
  __shared__ T sd[G::N1];
  __shared__ uint16_t sc[G::N1];
  __shared__ T s2[G::N1];  // first post-sweep
  const MgLevel& L = a.L;
  const MgLevel& Lc = a.Lc;
  const T h2[3] = {T(L.h2[0]), T(L.h2[1]), T(L.h2[2])};
  const int tid = threadIdx.x;
  const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY, k0 = blockIdx.z * TZ;
  const int sx = Lc.fx - 1, sy = Lc.fy - 1, sz = Lc.fz - 1;

  // x and e over tile + 2; b and the code over tile + 1
  T xv[G::S2];
  int r1[G::S2];
  uint16_t cv[G::S2];
  T bw[G::S2];
#pragma unroll
  for (int q = 0; q < G::S2; ++q) {
    const int s = tid + q * KT;
    r1[q] = -1;
    if (s < G::N2) {
      const int ai = s % G::RX2, t = s / G::RX2, bj = t % G::RY2, ck = t / G::RY2;
      const int gi = wrapc(i0 - 2 + ai, L.nx), gj = wrapc(j0 - 2 + bj, L.ny);
      const int gk = ZD ? wrapc(k0 - HZ + ck, L.nz) : 0;
      const int c = gi + L.nx * (gj + L.ny * gk);
      r1[q] = G::r1of(ai, bj, ck);
      const T x2 = BF ? T(__bfloat162float(a.xaB[c])) : a.xa[c];
      xv[q] = x2 + a.e[(gi >> sx) + Lc.nx * ((gj >> sy) + Lc.ny * (gk >> sz))];
      if (r1[q] >= 0) {
        cv[q] = L.code[c];
        bw[q] = a.b[c];
      }
    }
  }
  // the tile's own cells: index and, for the dots, r_new and r
  int oc[G::ST];
  double rnv[G::ST], rv[G::ST];
#pragma unroll
  for (int q = 0; q < G::ST; ++q) {
    const int s = tid + q * KT;
    oc[q] = -1;
    if (s >= G::NT) continue;
    const int ai = s % TX, t = s / TX, bj = t % TY, ck = t / TY;
    const int gi = i0 + ai, gj = j0 + bj, gk = k0 + ck;
    if (gi >= L.nx || gj >= L.ny || gk >= L.nz) continue;
    oc[q] = gi + L.nx * (gj + L.ny * gk);
    if constexpr (DST == UD_DOTS) {
      rnv[q] = a.rn[oc[q]];
      rv[q] = a.r[oc[q]];
    }
  }
  double mv[1];
  if (stopOrMeans<1>(a.ctl, {a.mb}, mv)) return;
  if (a.countIt && blk == 0 && tid == 0) stA(&a.ctl->it, ldA(&a.ctl->it) + 1);
  const T m = T(mv[0]);
  // ---- phase A: prolonged x (x + e of the parent) over tile + 2 ----
#pragma unroll
  for (int q = 0; q < G::S2; ++q) {
    const int s = tid + q * KT;
    if (s >= G::N2) continue;
    s1[s] = xv[q];
    if (r1[q] >= 0) {
      const uint16_t code = cv[q];
      T bv = bw[q];
      if (a.mb.on() && (code & kFluidBit)) bv = bv - m;
      sb[r1[q]] = bv;
      sd[r1[q]] = diagSum<T>(code, h2);
      sc[r1[q]] = code;
    }
  }
  __syncthreads();
  // ---- phase B: first post-sweep over tile + 1 ----
#pragma unroll
  for (int q = 0; q < G::S1; ++q) {
    const int s = tid + q * KT;
    if (s >= G::N1) continue;
    const int ai = s % G::RX1, t = s / G::RX1, bj = t % G::RY1, ck = t / G::RY1;
    const int i2 = (ai + 1) + G::RX2 * ((bj + 1) + G::RY2 * (ck + HZ / 2));
    T xn[6];
    nbrs<T, G::RX2, G::P2, ZD>(s1, i2, xn);
    s2[s] = sweepD<T>(sc[s], sd[s], s1[i2], xn, sb[s], h2);
  }
  __syncthreads();
  // ---- phase C: second post-sweep on the tile ----
  double acc[2] = {0.0, 0.0};
  T outv[G::ST];
#pragma unroll
  for (int q = 0; q < G::ST; ++q) {
    outv[q] = T(0);
    if (oc[q] < 0) continue;
    const int s = tid + q * KT;
    const int ai = s % TX, t = s / TX, bj = t % TY, ck = t / TY;
    const int i1 = (ai + 1) + G::RX1 * ((bj + 1) + G::RY1 * (ck + HZ / 2));
    const uint16_t code = sc[i1];
    T xn[6];
    nbrs<T, G::RX1, G::P1, ZD>(s2, i1, xn);
    outv[q] = sweepD<T>(code, sd[i1], s2[i1], xn, sb[i1], h2);
    if constexpr (DST != UD_X) {
      const double v = (code & kFluidBit) ? double(outv[q]) : 0.0;
      if constexpr (DST == UD_DOTS) {
        acc[0] += v * rnv[q];
        acc[1] += v * rv[q];
      } else {
        acc[0] += v;
      }
    }
  }
  if constexpr (DST == UD_DOTS) {
    blockToPartials<2>(acc, a.rb.partials, a.rb.stride, blk);
  } else if constexpr (DST == UD_Z) {
    if (a.rb.partials) {
      double one[1] = {acc[0]};
      blockToPartials<1>(one, a.rb.partials, a.rb.stride, blk);
    }
  }
  // ---- stores ----
#pragma unroll
  for (int q = 0; q < G::ST; ++q) {
    if (oc[q] < 0) continue;
    if constexpr (DST == UD_X) {
      a.xb[oc[q]] = outv[q];
    } else {
      // sweepD returns 0 on non-fluid cells, so double(out) is z's value there too
      a.z[oc[q]] = double(outv[q]);
    }
  }
}

// Tile shapes: 16 x 8 x 4 cells on 3-D levels, 16 x 16 on 2-D ones.
template <class F>
void withTile(const MgLevel& L, F&& f) {
  if (L.nz > 1) f(std::integral_constant<int, 16>{}, std::integral_constant<int, 8>{}, std::integral_constant<int, 4>{},
                  std::integral_constant<int, 2>{});
  else f(std::integral_constant<int, 16>{}, std::integral_constant<int, 16>{}, std::integral_constant<int, 1>{},
         std::integral_constant<int, 0>{});
}

}  // namespace

dim3 tileGrid(const MgLevel& L) {
  dim3 g;
  withTile(L, [&](auto tx, auto ty, auto tz, auto) {
    g = dim3((L.nx + tx() - 1) / tx(), (L.ny + ty() - 1) / ty(), (L.nz + tz() - 1) / tz());
  });
  return g;
}
int tileBlocks(const MgLevel& L) {
  const dim3 g = tileGrid(L);
  return int(g.x * g.y * g.z);
}
bool tileFits(const MgLevel& L, const MgLevel& Lc) {
  // the tile must cover whole coarse cells along every axis
  bool ok = true;
  withTile(L, [&](auto tx, auto ty, auto tz, auto) {
    ok = tx() % Lc.fx == 0 && ty() % Lc.fy == 0 && tz() % Lc.fz == 0;
  });
  return ok;
}

template <class T>
void launch_down(int src, const DownArgs<T>& a, hipStream_t s) {
  if (!tileFits(a.L, a.Lc)) throw std::logic_error("down kernel: tile does not cover whole coarse cells");
  const dim3 grid = tileGrid(a.L);
  withTile(a.L, [&](auto tx, auto ty, auto tz, auto hz) {
    constexpr int TX = decltype(tx)::value, TY = decltype(ty)::value, TZ = decltype(tz)::value,
                  HZ = decltype(hz)::value;
    if (a.xaB) {
      if constexpr (std::is_same_v<T, float>) {
        if (src == DS_B) k_down<T, DS_B, TX, TY, TZ, HZ, true><<<grid, KT, 0, s>>>(a);
        else if (src == DS_UPD) k_down<T, DS_UPD, TX, TY, TZ, HZ, true><<<grid, KT, 0, s>>>(a);
        else k_down<T, DS_RN, TX, TY, TZ, HZ, true><<<grid, KT, 0, s>>>(a);
      } else {
        throw std::logic_error("down kernel: BF16 scratch needs the FP32 cycle");
      }
    } else if (src == DS_B) {
      k_down<T, DS_B, TX, TY, TZ, HZ, false><<<grid, KT, 0, s>>>(a);
    } else if (src == DS_UPD) {
      k_down<T, DS_UPD, TX, TY, TZ, HZ, false><<<grid, KT, 0, s>>>(a);
    } else {
      k_down<T, DS_RN, TX, TY, TZ, HZ, false><<<grid, KT, 0, s>>>(a);
    }
  });
  NBL_LAUNCH_CHECK();
}

template <class T>
void launch_up(int dst, const UpArgs<T>& a, hipStream_t s) {
  const dim3 grid = tileGrid(a.L);
  withTile(a.L, [&](auto tx, auto ty, auto tz, auto hz) {
    constexpr int TX = decltype(tx)::value, TY = decltype(ty)::value, TZ = decltype(tz)::value,
                  HZ = decltype(hz)::value;
    if (a.xaB) {
      if constexpr (std::is_same_v<T, float>) {
        if (dst == UD_X) k_up<T, UD_X, TX, TY, TZ, HZ, true><<<grid, KT, 0, s>>>(a);
        else if (dst == UD_Z) k_up<T, UD_Z, TX, TY, TZ, HZ, true><<<grid, KT, 0, s>>>(a);
        else k_up<T, UD_DOTS, TX, TY, TZ, HZ, true><<<grid, KT, 0, s>>>(a);
      } else {
        throw std::logic_error("up kernel: BF16 scratch needs the FP32 cycle");
      }
    } else if (dst == UD_X) {
      k_up<T, UD_X, TX, TY, TZ, HZ, false><<<grid, KT, 0, s>>>(a);
    } else if (dst == UD_Z) {
      k_up<T, UD_Z, TX, TY, TZ, HZ, false><<<grid, KT, 0, s>>>(a);
    } else {
      k_up<T, UD_DOTS, TX, TY, TZ, HZ, false><<<grid, KT, 0, s>>>(a);
    }
  });
  NBL_LAUNCH_CHECK();
}

template void launch_down<double>(int, const DownArgs<double>&, hipStream_t);
template void launch_down<float>(int, const DownArgs<float>&, hipStream_t);
template void launch_up<double>(int, const UpArgs<double>&, hipStream_t);
template void launch_up<float>(int, const UpArgs<float>&, hipStream_t);

}  // namespace nbl
