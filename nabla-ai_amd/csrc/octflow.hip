// octflow.hip — flow step on octree leaves (collocated finite volumes).
//
// No reference implementation exists: the reference's design record plans an
// octree-native collocated finite-volume solve (docs/adr/0008-octree-native-
// flow-solve.md:35-86) and implements its first brick, the OctreePoisson
// operator (octree_poisson.cpp:64-105), which the projection here solves with
// the octree multigrid CG.  The time integration is the MAC solver's SSP-RK3
// with a projection after every stage (ns_solver.cpp:773-795), the step size
// its CFL rule per leaf (ns_solver.cpp:581-603).
//
// One leaf per thread in Morton order.  Faces are walked as the face table
// holds them (direction -x,+x,-y,+y,-z,+z, then neighbour order); an interior
// face's area is the finer leaf's face (the transport step's rule,
// transport.cpp:41-45) and its distance the centre-to-centre one along the
// axis (the OctreePoisson coefficient).  Every interior face term is formed
// from both sides with the same operands in the same order and opposite sign,
// so fluxes cancel exactly across every face, level jumps included.  Face
// velocities U (outward from the leaf that stores them) live per face entry:
// the first neighbour of direction d of leaf i at d*n + i (which also holds a
// domain-boundary face's value), further neighbours at 6n + their xnb index.
#include <cstring>
#include <stdexcept>

#include "dev.hpp"
#include "octflow.hpp"
#include "weno.hpp"

namespace nbl {
namespace {

constexpr int kFlowBlocks = 4096;  // partial slots of the reducing kernels
int fgrid(long long n) { return gridFor(n, kBlock, kFlowBlocks); }

#define FGRID(i, n) for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

struct TreeV {
  long long n;
  Geo G;
  const uint64_t* __restrict__ keys;
  const uint8_t* __restrict__ lev;
  const uint8_t* __restrict__ mask;
  FaceWalk F;
};

struct Bc {
  int face[6];
  double uin;  // mean inflow speed along the inward normal
  int profile;  // 0 uniform, 1 plane Poiseuille across the next axis
  double nu;
  int weno;
};

// Inflow speed at fluid leaf i's domain face d.
__device__ __forceinline__ double inletSpeed(const TreeV& T, const Bc& B, long long i, int d) {
  if (B.profile == 0) return B.uin;
  const int ax = ((d >> 1) + 1) % 3;
  const double len = T.G.fine[ax] * double(1u << T.G.maxL);
  const double s = (centerAxis(T.G, T.keys[i], T.lev[i], ax) - T.G.org[ax]) / len;
  return 6.0 * B.uin * s * (1.0 - s);
}

__device__ __forceinline__ double faceArea(const Geo& G, int lvl, int axis) {
  return sizeAxis(G, lvl, (axis + 1) % 3) * sizeAxis(G, lvl, (axis + 2) % 3);
}
// the neighbour across direction d when it is one fluid leaf of the same level
__device__ __forceinline__ int sameLevel(const TreeV& T, int c, int d) {
  if (c < 0 || T.F.cnt[6 * (long long)c + d] != 1) return -1;
  const int nb = T.F.first[(long long)d * T.n + c];
  return (T.lev[nb] == T.lev[c] && T.mask[nb] == 0) ? nb : -1;
}

// WENO-5 line of every first-neighbour face entry (d, i) whose neighbour is one
// fluid leaf of i's level: (m2, m1, p1, p2), the two same-level fluid leaves
// behind i and beyond the neighbour along the face normal, or -1s where the
// line is incomplete (k_of_rhs then uses the upwind value).  The tree and its
// masks are fixed for a flow's life, so the table is built once; it replaces
// four dependent neighbour-of-neighbour lookups per face in every stage.
__global__ void k_of_wenoline(TreeV T, int4* __restrict__ wl) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * T.n) return;
  const int d = int(t / T.n);
  const long long i = t - (long long)d * T.n;
  int4 w = make_int4(-1, -1, -1, -1);
  if (T.mask[i] == 0 && T.F.cnt[6 * i + d] == 1) {
    const int nb = T.F.first[(long long)d * T.n + i];
    if (T.mask[nb] == 0 && T.lev[nb] == T.lev[i]) {
      const int od = d ^ 1;
      const int m1 = sameLevel(T, int(i), od);
      const int m2 = sameLevel(T, m1, od);
      const int p1 = sameLevel(T, nb, d);
      const int p2 = sameLevel(T, p1, d);
      if (m2 >= 0 && p2 >= 0 && m1 >= 0 && p1 >= 0) w = make_int4(m2, m1, p1, p2);
    }
  }
  wl[t] = w;
}

// Walks leaf i's faces: f(d, nb, entry) with nb < 0 for a domain-boundary face.
template <class Fn>
__device__ __forceinline__ void eachFace(const TreeV& T, long long i, Fn&& f) {
  int x = T.F.xrow[i];
  // unrolled: d indexes the kernel arguments' per-face and per-axis arrays,
  // which a runtime index would copy into scratch memory
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    const int c = T.F.cnt[6 * i + d];
    if (c == 0) {
      f(d, -1, (long long)d * T.n + i);
      continue;
    }
    for (int m = 0; m < c; ++m) {
      if (m == 0) {
        f(d, T.F.first[(long long)d * T.n + i], (long long)d * T.n + i);
      } else {
        f(d, T.F.xnb[x], 6 * T.n + x);
        ++x;
      }
    }
  }
}

// Boundary volume fluxes over fluid leaves: inflow, outflow of the cell
// velocities, outlet area.
__global__ void __launch_bounds__(kBlock) k_of_bflux(TreeV T, Bc B, const double* __restrict__ u,
                                                     const double* __restrict__ v, const double* __restrict__ w,
                                                     RedBuf rb) {
  double acc[3] = {0.0, 0.0, 0.0};
  FGRID(i, T.n) {
    if (T.mask[i] != 0) continue;
    const int li = T.lev[i];
    for (int d = 0; d < 6; ++d) {
      if (T.F.cnt[6 * i + d] != 0) continue;
      const int axis = d >> 1;
      const double sign = (d & 1) ? 1.0 : -1.0;
      const double A = faceArea(T.G, li, axis);
      if (B.face[d] == NBL_OCT_INLET) acc[0] += inletSpeed(T, B, i, d) * A;
      if (B.face[d] == NBL_OCT_OUTLET) {
        const double vel = axis == 0 ? u[i] : axis == 1 ? v[i] : w[i];
        acc[1] += sign * vel * A;
        acc[2] += A;
      }
    }
  }
  blockToPartials<3>(acc, rb.partials, rb.stride, blockIdx.x);
}

// scal[0] = uniform outlet velocity correction c (outflow + c * area = inflow),
// scal[1] = inflow, scal[2] = corrected outflow.
__global__ void __launch_bounds__(kBlock) k_of_corr(RedBuf rb, int parts, double* scal) {
  double t[3];
  foldPartials<3>(rb.partials, rb.stride, parts, t);
  if (threadIdx.x == 0) {
    const double c = t[2] > 0.0 ? (t[0] - t[1]) / t[2] : 0.0;
    scal[0] = c;
    scal[1] = t[0];
    scal[2] = t[1] + c * t[2];
  }
}

// outward normal velocity of a domain-boundary face d of fluid leaf i
__device__ __forceinline__ double boundaryU(const TreeV& T, const Bc& B, long long i, int d, double velAxis,
                                            double c) {
  const double sign = (d & 1) ? 1.0 : -1.0;
  switch (B.face[d]) {
    case NBL_OCT_INLET: return -inletSpeed(T, B, i, d);
    case NBL_OCT_OUTLET: return sign * velAxis + c;
    default: return 0.0;
  }
}

// Poisson right-hand side -(sum_f U~_f A_f)/dt of the provisional face
// velocities U~_f = 0.5 (u_P + u_N).n (fluid leaves; 0 elsewhere).
__global__ void __launch_bounds__(kBlock) k_of_div(TreeV T, Bc B, const double* __restrict__ u,
                                                   const double* __restrict__ v, const double* __restrict__ w,
                                                   const double* __restrict__ scal, double dt,
                                                   double* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  if (T.mask[i] != 0) {
    b[i] = 0.0;
    return;
  }
  const double c0 = scal[0];
  const int li = T.lev[i];
  double D = 0.0;
  eachFace(T, i, [&](int d, int nb, long long) {
    const int axis = d >> 1;
    const double sign = (d & 1) ? 1.0 : -1.0;
    const double* __restrict__ vel = axis == 0 ? u : axis == 1 ? v : w;
    if (nb < 0) {
      D += boundaryU(T, B, i, d, vel[i], c0) * faceArea(T.G, li, axis);
    } else if (T.mask[nb] == 0) {
      const int ln = T.lev[nb];
      D += sign * 0.5 * (vel[i] + vel[nb]) * faceArea(T.G, ln > li ? ln : li, axis);
    }
  });
  b[i] = -D / dt;
}

// Projection: U_f = U~_f - dt (p_N - p_P)/dist_f on fluid-fluid faces,
// boundary values elsewhere (0 at non-fluid neighbours), stored per entry;
// the cell velocity minus dt times the Gauss gradient of p (face value the
// mean of the two leaves, the leaf's own across boundary and non-fluid faces)
// into uo; partials: max |sum_f U_f A_f| / V and sum 0.5 |u|^2 V.
__global__ void __launch_bounds__(kBlock) k_of_correct(TreeV T, Bc B, const double* __restrict__ u,
                                                       const double* __restrict__ v, const double* __restrict__ w,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ scal, double dt,
                                                       double* __restrict__ U, double* __restrict__ uo,
                                                       double* __restrict__ vo, double* __restrict__ wo, RedBuf rb) {
  double acc[2] = {0.0, 0.0};
  FGRID(i, T.n) {
    if (T.mask[i] != 0) {
      eachFace(T, i, [&](int, int, long long e) { U[e] = 0.0; });
      uo[i] = 0.0;
      vo[i] = 0.0;
      wo[i] = 0.0;
      continue;
    }
    const double c0 = scal[0];
    const int li = T.lev[i];
    const uint64_t ki = T.keys[i];
    const double pi = p[i];
    // the gradient's components as scalars: an array indexed by the face's
    // axis would live in scratch memory
    double div = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    eachFace(T, i, [&](int d, int nb, long long e) {
      const int axis = d >> 1;
      const double sign = (d & 1) ? 1.0 : -1.0;
      const double* __restrict__ vel = axis == 0 ? u : axis == 1 ? v : w;
      double Uf;
      if (nb < 0) {
        Uf = boundaryU(T, B, i, d, vel[i], c0);
        div += Uf * faceArea(T.G, li, axis);
      } else if (T.mask[nb] != 0) {
        Uf = 0.0;
      } else {
        const int ln = T.lev[nb];
        const double A = faceArea(T.G, ln > li ? ln : li, axis);
        const double dist = fabs(centerAxis(T.G, T.keys[nb], ln, axis) - centerAxis(T.G, ki, li, axis));
        const double dp = p[nb] - pi;
        // N lies along the outward normal: dp / dist is the normal derivative
        Uf = sign * 0.5 * (vel[i] + vel[nb]) - dt * dp / dist;
        div += Uf * A;
        const double ga = sign * (0.5 * dp) * A;
        if (axis == 0) gx += ga;
        else if (axis == 1) gy += ga;
        else gz += ga;
      }
      U[e] = Uf;
    });
    const double V = sizeAxis(T.G, li, 0) * sizeAxis(T.G, li, 1) * sizeAxis(T.G, li, 2);
    const double un = u[i] - dt * gx / V, vn = v[i] - dt * gy / V, wn = w[i] - dt * gz / V;
    uo[i] = un;
    vo[i] = vn;
    wo[i] = wn;
    acc[0] = fmax(acc[0], fabs(div) / V);
    acc[1] += 0.5 * (un * un + vn * vn + wn * wn) * V;
  }
  blockToPartials<2, 1u>(acc, rb.partials, rb.stride, blockIdx.x);
}

// Face velocities of a given cell field (no pressure): U = 0.5 (u_P + u_N).n,
// boundary values with no outlet correction, 0 at non-fluid faces.
__global__ void __launch_bounds__(kBlock) k_of_faceinit(TreeV T, Bc B, const double* __restrict__ u,
                                                        const double* __restrict__ v,
                                                        const double* __restrict__ w, double* __restrict__ U) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  const bool fl = T.mask[i] == 0;
  eachFace(T, i, [&](int d, int nb, long long e) {
    const int axis = d >> 1;
    const double sign = (d & 1) ? 1.0 : -1.0;
    const double* __restrict__ vel = axis == 0 ? u : axis == 1 ? v : w;
    double Uf = 0.0;
    if (fl && nb < 0) Uf = boundaryU(T, B, i, d, vel[i], 0.0);
    else if (fl && T.mask[nb] == 0) Uf = sign * 0.5 * (vel[i] + vel[nb]);
    U[e] = Uf;
  });
}

// sum_f U_f A_f / V of the stored face velocities (fluid leaves; 0 elsewhere).
__global__ void __launch_bounds__(kBlock) k_of_divergence(TreeV T, const double* __restrict__ U,
                                                          double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  if (T.mask[i] != 0) {
    out[i] = 0.0;
    return;
  }
  const int li = T.lev[i];
  double div = 0.0;
  eachFace(T, i, [&](int d, int nb, long long e) {
    const int axis = d >> 1;
    if (nb >= 0 && T.mask[nb] != 0) return;
    const int lf = nb >= 0 && T.lev[nb] > li ? T.lev[nb] : li;
    div += U[e] * faceArea(T.G, lf, axis);
  });
  out[i] = div / (sizeAxis(T.G, li, 0) * sizeAxis(T.G, li, 1) * sizeAxis(T.G, li, 2));
}

// Momentum right-hand side of fluid leaves (0 elsewhere):
// (-sum_f U_f A_f phi_f + nu sum_f A_f (phi_N - phi_P)/dist_f + boundary
// terms) / V per component.  phi_f: WENO-5 along the face's five-leaf line
// of one level (two leaves upstream, two downstream, all fluid), else the
// upwind leaf's value.  Non-fluid neighbours and walls: no-slip (phi = 0 at
// the face, half a leaf away); free-slip walls: the normal component only;
// inlet: the inflow vector; outlet: zero gradient.  The RK stage is applied
// in the same pass: out = a u0 + b (u + dt R) on fluid leaves, 0 elsewhere
// (u itself is read at neighbours, so the result goes to a second buffer).
__global__ void __launch_bounds__(kBlock) k_of_rhs(TreeV T, Bc B, const double* __restrict__ u,
                                                   const double* __restrict__ v, const double* __restrict__ w,
                                                   const double* __restrict__ U, const double* __restrict__ u0,
                                                   const double* __restrict__ v0, const double* __restrict__ w0,
                                                   double a, double b, double dt, double* __restrict__ ou,
                                                   double* __restrict__ ov, double* __restrict__ ow,
                                                   const int4* __restrict__ wl) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  if (T.mask[i] != 0) {
    ou[i] = 0.0;
    ov[i] = 0.0;
    ow[i] = 0.0;
    return;
  }
  const double* __restrict__ ph[3] = {u, v, w};
  const int li = T.lev[i];
  const uint64_t ki = T.keys[i];
  const double hi[3] = {sizeAxis(T.G, li, 0), sizeAxis(T.G, li, 1), sizeAxis(T.G, li, 2)};
  const double pP[3] = {u[i], v[i], w[i]};
  double conv[3] = {0.0, 0.0, 0.0}, diff[3] = {0.0, 0.0, 0.0};
  eachFace(T, i, [&](int d, int nb, long long e) {
    const int axis = d >> 1;
    const double sign = (d & 1) ? 1.0 : -1.0;
    if (nb < 0) {
      const double A = faceArea(T.G, li, axis);
      const double half = 0.5 * hi[axis];
      switch (B.face[d]) {
        case NBL_OCT_WALL:
          for (int c = 0; c < 3; ++c) diff[c] += B.nu * A * (0.0 - pP[c]) / half;
          break;
        case NBL_OCT_SLIP:
          diff[axis] += B.nu * A * (0.0 - pP[axis]) / half;
          break;
        case NBL_OCT_INLET: {
          const double Ub = U[e];
          const double sp = inletSpeed(T, B, i, d);
          for (int c = 0; c < 3; ++c) {
            const double pin = c == axis ? -sign * sp : 0.0;
            conv[c] -= Ub * A * pin;
            diff[c] += B.nu * A * (pin - pP[c]) / half;
          }
          break;
        }
        default: {  // outlet
          const double Ub = U[e];
          for (int c = 0; c < 3; ++c) conv[c] -= Ub * A * pP[c];
          break;
        }
      }
      return;
    }
    const int ln = T.lev[nb];
    const double A = faceArea(T.G, ln > li ? ln : li, axis);
    if (T.mask[nb] != 0) {  // no-slip at the non-fluid leaf's face
      const double half = 0.5 * hi[axis];
      for (int c = 0; c < 3; ++c) diff[c] += B.nu * A * (0.0 - pP[c]) / half;
      return;
    }
    const double dist = fabs(centerAxis(T.G, T.keys[nb], ln, axis) - centerAxis(T.G, ki, li, axis));
    const double Uf = U[e];
    int m1 = -1, m2 = -1, p1 = -1, p2 = -1;
    if (B.weno && ln == li && T.F.cnt[6 * i + d] == 1) {
      const int4 w = wl[(long long)d * T.n + i];  // (k_of_wenoline)
      m2 = w.x;
      m1 = w.y;
      p1 = w.z;
      p2 = w.w;
    }
    const bool line = m2 >= 0 && p2 >= 0 && m1 >= 0 && p1 >= 0;
    for (int c = 0; c < 3; ++c) {
      const double* __restrict__ q = ph[c];
      const double qN = q[nb];
      double qf;
      if (line) qf = Uf >= 0.0 ? weno5(q[m2], q[m1], pP[c], qN, q[p1]) : weno5(q[p2], q[p1], qN, pP[c], q[m1]);
      else qf = Uf >= 0.0 ? pP[c] : qN;
      conv[c] -= Uf * A * qf;
      diff[c] += B.nu * A * (qN - pP[c]) / dist;
    }
  });
  const double V = hi[0] * hi[1] * hi[2];
  ou[i] = a * u0[i] + b * (pP[0] + dt * ((conv[0] + diff[0]) / V));
  ov[i] = a * v0[i] + b * (pP[1] + dt * ((conv[1] + diff[1]) / V));
  ow[i] = a * w0[i] + b * (pP[2] + dt * ((conv[2] + diff[2]) / V));
}

// u = v = w = 0 on non-fluid leaves.
__global__ void k_of_solid_zero(long long n, const uint8_t* __restrict__ mask, double* __restrict__ u,
                                double* __restrict__ v, double* __restrict__ w) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || mask[i] == 0) return;
  u[i] = 0.0;
  v[i] = 0.0;
  w[i] = 0.0;
}

// partials of max over fluid leaves of |u|/hx + |v|/hy + |w|/hz + 2 nu sum 1/h^2
__global__ void __launch_bounds__(kBlock) k_of_cfl(TreeV T, double nu, const double* __restrict__ u,
                                                   const double* __restrict__ v, const double* __restrict__ w,
                                                   RedBuf rb) {
  double acc[1] = {0.0};
  FGRID(i, T.n) {
    if (T.mask[i] != 0) continue;
    const int li = T.lev[i];
    const double hx = sizeAxis(T.G, li, 0), hy = sizeAxis(T.G, li, 1), hz = sizeAxis(T.G, li, 2);
    const double visc = 2.0 * nu * (1.0 / (hx * hx) + 1.0 / (hy * hy) + 1.0 / (hz * hz));
    const double inv = fabs(u[i]) / hx + fabs(v[i]) / hy + fabs(w[i]) / hz + visc;
    acc[0] = fmax(acc[0], inv);
  }
  blockToPartials<1, 1u>(acc, rb.partials, rb.stride, blockIdx.x);
}

template <int K, unsigned MAXMASK>
__global__ void __launch_bounds__(kBlock) k_of_fold(RedBuf rb, int parts, double* out) {
  double t[K];
  foldPartials<K, MAXMASK>(rb.partials, rb.stride, parts, t);
  if (threadIdx.x == 0)
    for (int q = 0; q < K; ++q) out[q] = t[q];
}

}  // namespace

namespace {
TreeV treeView(const OctreeDev& t) {
  const OctLevel& L = t.level0();
  return TreeV{(long long)L.n, t.geo(), L.keys.as<uint64_t>(), L.lev.as<uint8_t>(), L.mask.as<uint8_t>(),
               t.faceWalk()};
}
Bc bcOf(const nbl_octree_flow_cfg& c) {
  Bc b;
  for (int d = 0; d < 6; ++d) b.face[d] = c.face[d];
  b.uin = c.inlet_speed;
  b.profile = c.inlet_profile;
  b.nu = c.nu;
  b.weno = c.convection;
  return b;
}
}  // namespace

OctFlowDev::OctFlowDev(nbl_ctx* ctx, OctreeDev& tree, const nbl_octree_flow_cfg& cfg)
    : ctx_(ctx), tree_(tree), cfg_(cfg), s_(tree.stream()), n_(tree.size()) {
  if (!(cfg.nu >= 0.0)) throw std::invalid_argument("octree flow: nu must be >= 0");
  if (!(cfg.cfl > 0.0)) throw std::invalid_argument("octree flow: cfl must be > 0");
  if (cfg.convection != 0 && cfg.convection != 1) throw std::invalid_argument("octree flow: convection 0 or 1");
  if (cfg.inlet_profile != 0 && cfg.inlet_profile != 1) throw std::invalid_argument("octree flow: inlet_profile 0 or 1");
  if (cfg.poisson_max_iters < 1 || !(cfg.poisson_tol > 0.0))
    throw std::invalid_argument("octree flow: poisson tolerance > 0 and iterations >= 1");
  for (int d = 0; d < 6; ++d)
    if (cfg.face[d] < NBL_OCT_WALL || cfg.face[d] > NBL_OCT_OUTLET)
      throw std::invalid_argument("octree flow: face type must be NBL_OCT_WALL, _SLIP, _INLET or _OUTLET");
  tree_.setMgPrecision(cfg.precision);
  for (int c = 0; c < 3; ++c) {
    u_[c] = DevBuf(ctx, n_ * 8);
    u0_[c] = DevBuf(ctx, n_ * 8);
    rhs_[c] = DevBuf(ctx, n_ * 8);
    NBL_HIP(hipMemsetAsync(u_[c].as<void>(), 0, n_ * 8, s_));
  }
  p_ = DevBuf(ctx, n_ * 8);
  b_ = DevBuf(ctx, n_ * 8);
  U_ = DevBuf(ctx, tree.faceEntries() * 8);
  red_ = DevBuf(ctx, size_t(3) * kFlowBlocks * 8);
  scal_ = DevBuf(ctx, 16 * 8);
  NBL_HIP(hipMemsetAsync(p_.as<void>(), 0, n_ * 8, s_));
  NBL_HIP(hipMemsetAsync(scal_.as<void>(), 0, 16 * 8, s_));
  NBL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hscal_), 16 * 8, hipHostMallocDefault));
  std::memset(hscal_, 0, 16 * 8);
  if (bcOf(cfg_).weno) {  // WENO-5 lines of the fixed tree
    wl_ = DevBuf(ctx, 6 * n_ * sizeof(int4));
    const TreeV T = treeView(tree_);
    k_of_wenoline<<<gridFor(6 * T.n), kBlock, 0, s_>>>(T, wl_.as<int4>());
    NBL_LAUNCH_CHECK();
  }
  faceInit();
  NBL_HIP(hipStreamSynchronize(s_));
}

OctFlowDev::~OctFlowDev() {
  if (hscal_) (void)hipHostFree(hscal_);
}

void OctFlowDev::faceInit() {
  const TreeV T = treeView(tree_);
  k_of_faceinit<<<gridFor(T.n), kBlock, 0, s_>>>(T, bcOf(cfg_), u_[0].as<double>(), u_[1].as<double>(),
                                                 u_[2].as<double>(), U_.as<double>());
  NBL_LAUNCH_CHECK();
}

void OctFlowDev::setState(const double* u, const double* v, const double* w, const double* p) {
  const double* src[3] = {u, v, w};
  for (int c = 0; c < 3; ++c) tree_.toSorted(src[c], u_[c].as<double>());
  if (p) tree_.toSorted(p, p_.as<double>());
  else NBL_HIP(hipMemsetAsync(p_.as<void>(), 0, n_ * 8, s_));
  // non-fluid leaves carry no velocity
  const long long N = (long long)n_;
  k_of_solid_zero<<<gridFor(N), kBlock, 0, s_>>>(N, tree_.level0().mask.as<uint8_t>(), u_[0].as<double>(),
                                                 u_[1].as<double>(), u_[2].as<double>());
  NBL_LAUNCH_CHECK();
  faceInit();
  NBL_HIP(hipStreamSynchronize(s_));
}

void OctFlowDev::getState(double* u, double* v, double* w, double* p) {
  double* dst[3] = {u, v, w};
  for (int c = 0; c < 3; ++c)
    if (dst[c]) tree_.fromSorted(u_[c].as<double>(), dst[c]);
  if (p) tree_.fromSorted(p_.as<double>(), p);
}

void OctFlowDev::divergence(double* out) {
  const TreeV T = treeView(tree_);
  k_of_divergence<<<gridFor(T.n), kBlock, 0, s_>>>(T, U_.as<double>(), b_.as<double>());
  NBL_LAUNCH_CHECK();
  tree_.fromSorted(b_.as<double>(), out);
}

double OctFlowDev::cflDt() {
  const TreeV T = treeView(tree_);
  const RedBuf rb{red_.as<double>(), kFlowBlocks};
  const int g = fgrid(T.n);
  k_of_cfl<<<g, kBlock, 0, s_>>>(T, cfg_.nu, u_[0].as<double>(), u_[1].as<double>(), u_[2].as<double>(), rb);
  k_of_fold<1, 1u><<<1, kBlock, 0, s_>>>(rb, g, scal_.as<double>() + 8);
  NBL_LAUNCH_CHECK();
  NBL_HIP(hipMemcpyAsync(hscal_, scal_.as<void>(), 16 * 8, hipMemcpyDeviceToHost, s_));
  NBL_HIP(hipStreamSynchronize(s_));
  const double m = hscal_[8];
  if (!(m > 0.0)) throw std::runtime_error("octree flow: no fluid leaf bounds the time step");
  return cfg_.cfl / m;
}

void OctFlowDev::project(double dt, int* it, double* resid, double* ms) {
  const TreeV T = treeView(tree_);
  const Bc B = bcOf(cfg_);
  const RedBuf rb{red_.as<double>(), kFlowBlocks};
  const int g = fgrid(T.n);
  double* sc = scal_.as<double>();
  k_of_bflux<<<g, kBlock, 0, s_>>>(T, B, u_[0].as<double>(), u_[1].as<double>(), u_[2].as<double>(), rb);
  k_of_corr<<<1, kBlock, 0, s_>>>(rb, g, sc);
  k_of_div<<<gridFor(T.n), kBlock, 0, s_>>>(T, B, u_[0].as<double>(), u_[1].as<double>(), u_[2].as<double>(), sc,
                                            dt, b_.as<double>());
  NBL_LAUNCH_CHECK();
  const nbl_poisson_report r = tree_.solveMgSorted(b_.as<double>(), p_.as<double>(), cfg_.poisson_tol,
                                                   cfg_.poisson_max_iters, s_, p_.as<double>());
  if (!r.converged) throw std::runtime_error("octree flow: pressure solve did not converge");
  *it += r.iterations;
  *resid = std::max(*resid, r.residual);
  *ms += tree_.lastSolveMs();
  k_of_correct<<<g, kBlock, 0, s_>>>(T, B, u_[0].as<double>(), u_[1].as<double>(), u_[2].as<double>(),
                                     p_.as<double>(), sc, dt, U_.as<double>(), rhs_[0].as<double>(),
                                     rhs_[1].as<double>(), rhs_[2].as<double>(), rb);
  k_of_fold<2, 1u><<<1, kBlock, 0, s_>>>(rb, g, sc + 3);
  NBL_LAUNCH_CHECK();
  for (int c = 0; c < 3; ++c) std::swap(u_[c], rhs_[c]);
}

nbl_octree_step_report OctFlowDev::step(double dtIn) {
  const double dt = dtIn > 0.0 ? dtIn : cflDt();
  const TreeV T = treeView(tree_);
  const Bc B = bcOf(cfg_);
  const long long N = T.n;
  for (int c = 0; c < 3; ++c)
    NBL_HIP(hipMemcpyAsync(u0_[c].as<void>(), u_[c].as<void>(), n_ * 8, hipMemcpyDeviceToDevice, s_));
  const double ca[3] = {0.0, 0.75, 1.0 / 3.0}, cb[3] = {1.0, 0.25, 2.0 / 3.0};  // ns_solver.cpp:773-795
  int it = 0;
  double resid = 0.0, ms = 0.0;
  for (int st = 0; st < 3; ++st) {
    k_of_rhs<<<gridFor(N), kBlock, 0, s_>>>(T, B, u_[0].as<double>(), u_[1].as<double>(), u_[2].as<double>(),
                                            U_.as<double>(), u0_[0].as<double>(), u0_[1].as<double>(),
                                            u0_[2].as<double>(), ca[st], cb[st], dt, rhs_[0].as<double>(),
                                            rhs_[1].as<double>(), rhs_[2].as<double>(), wl_.as<int4>());
    NBL_LAUNCH_CHECK();
    for (int c = 0; c < 3; ++c) std::swap(u_[c], rhs_[c]);
    // The stage's velocity carries its Euler step with weight cb, so it is
    // projected with the step cb dt: p is then the pressure at every stage
    // (the same field up to the time the stage lands on), and each solve
    // starts from the last stage's p.
    project(cb[st] * dt, &it, &resid, &ms);
  }
  NBL_HIP(hipMemcpyAsync(hscal_, scal_.as<void>(), 16 * 8, hipMemcpyDeviceToHost, s_));
  NBL_HIP(hipStreamSynchronize(s_));
  ++step_;
  t_ += dt;
  nbl_octree_step_report rep{};
  rep.step = step_;
  rep.t = t_;
  rep.dt = dt;
  rep.inflow = hscal_[1];
  rep.outflow = hscal_[2];
  rep.divergence_max = hscal_[3];
  rep.kinetic_energy = hscal_[4];
  rep.poisson_iters = it;
  rep.poisson_residual = resid;
  rep.solve_ms = ms;
  return rep;
}

}  // namespace nbl
