// ns.hip — gfx950 kernels of the MAC time step (reference core/src/flow/ns_solver.cpp).
//
// Layout in HBM: the reference's own face arrays (u: (nx+1)*ny*nz, v:
// nx*(ny+1)*nz, w: nx*ny*(nz+1), compact index i fastest) and compact cell
// arrays.  Every kernel is a streaming pass over one flattened index space
// (faces of u|v|w, optionally followed by cells), one element per thread, so
// loads along i are coalesced.  Arithmetic follows the reference expression
// order, and the library is built with -ffp-contract=off, so per-element
// results are bit-identical to the reference; only reductions (CFL max,
// diagnostics sums) use a different (deterministic, tree) order.
#include <algorithm>
#include <stdexcept>

#include "dev.hpp"
#include "nbl_api.h"
#include "weno.hpp"

namespace nbl {
namespace {

struct Idx {  // staggered index helpers (ns_solver.hpp:93-110); arrays < 2^31 elements
  int nx, ny, nz;
  __device__ int u(int i, int j, int k) const { return i + (nx + 1) * (j + ny * k); }
  __device__ int v(int i, int j, int k) const { return i + nx * (j + (ny + 1) * k); }
  __device__ int w(int i, int j, int k) const { return i + nx * (j + ny * k); }
  __device__ int c(int i, int j, int k) const { return i + nx * (j + ny * k); }
};

__device__ __forceinline__ int wrapi(int a, int n) { return (a % n + n) % n; }

__device__ __forceinline__ bool solidAt(const DevGrid& g, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return false;
  return g.solid[i + g.nx * (j + g.ny * k)] != 0;
}

// Inlet velocity component `comp` on face f at transverse index a (used as the
// y index by the parabolic profile), ns_solver.cpp:71-93.
__device__ double inletComp(const DevGrid& g, int f, int a, int comp) {
  double mag = g.inletSpeed[f];
  if (g.profile[f] == NBL_INLET_PARABOLIC) {
    const double H = g.ny * g.dy;
    const double yc = g.oy + 0.5 * H;
    const double y = g.oy + (a + 0.5) * g.dy;
    const double r = (y - yc) / (0.5 * H);
    mag = 1.5 * g.inletSpeed[f] * (1.0 - r * r);
  }
  const int ax = g.inletAxis[f] == 0 ? 0 : (g.inletAxis[f] == 1 ? 1 : 2);
  return ax == comp ? mag : 0.0;
}

// Reflections pending while a ghost index is resolved: the reference evaluates
// `2*X - inner` recursively, so they are applied innermost first.
struct Refl {
  double X0 = 0.0, X1 = 0.0, X2 = 0.0;
  int d = 0;
  __device__ __forceinline__ bool push(double x) {
    if (d == 0) X0 = x;
    else if (d == 1) X1 = x;
    else if (d == 2) X2 = x;
    else return false;
    ++d;
    return true;
  }
  __device__ __forceinline__ double fold(double val) const {
    if (d >= 3) val = 2.0 * X2 - val;
    if (d >= 2) val = 2.0 * X1 - val;
    if (d >= 1) val = 2.0 * X0 - val;
    return val;
  }
};

// Ghost-aware reads (ns_solver.cpp:116-219), resolved iteratively.
__device__ double ghostU(const DevGrid& g, const double* __restrict__ u, int i, int j, int k) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Idx I{nx, ny, nz};
  Refl R;
  for (int guard = 0; guard < 12; ++guard) {
    if (k < 0 || k >= nz) {
      if (g.bc[4] == NBL_BC_PERIODIC) { k = wrapi(k, nz); continue; }
      if (!R.push(g.wall[k < 0 ? 4 : 5][0])) break;
      k = k < 0 ? -1 - k : 2 * nz - 1 - k;
      continue;
    }
    if (j < 0 || j >= ny) {
      if (g.bc[2] == NBL_BC_PERIODIC) { j = wrapi(j, ny); continue; }
      if (!R.push(g.wall[j < 0 ? 2 : 3][0])) break;
      j = j < 0 ? -1 - j : 2 * ny - 1 - j;
      continue;
    }
    if (i < 0 || i > nx) {
      if (g.bc[0] == NBL_BC_PERIODIC) { i = wrapi(i, nx); continue; }
      const int ib = i < 0 ? 0 : nx;
      if (g.bc[i < 0 ? 0 : 1] == NBL_BC_OUTLET) return R.fold(u[I.u(ib, j, k)]);
      if (!R.push(u[I.u(ib, j, k)])) break;
      i = 2 * ib - i;
      continue;
    }
    return R.fold(u[I.u(i, j, k)]);
  }
  return 0.0;
}

__device__ double ghostV(const DevGrid& g, const double* __restrict__ v, int i, int j, int k) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Idx I{nx, ny, nz};
  Refl R;
  for (int guard = 0; guard < 12; ++guard) {
    if (k < 0 || k >= nz) {
      if (g.bc[4] == NBL_BC_PERIODIC) { k = wrapi(k, nz); continue; }
      if (!R.push(g.wall[k < 0 ? 4 : 5][1])) break;
      k = k < 0 ? -1 - k : 2 * nz - 1 - k;
      continue;
    }
    if (i < 0 || i >= nx) {
      if (g.bc[0] == NBL_BC_PERIODIC) { i = wrapi(i, nx); continue; }
      const int f = i < 0 ? 0 : 1;
      const int im = i < 0 ? -1 - i : 2 * nx - 1 - i;
      if (g.bc[f] == NBL_BC_INLET) {
        if (!R.push(inletComp(g, f, j, 1))) break;
        i = im;
        continue;
      }
      if (g.bc[f] == NBL_BC_OUTLET) return R.fold(v[I.v(i < 0 ? 0 : nx - 1, j, k)]);
      if (!R.push(g.wall[f][1])) break;
      i = im;
      continue;
    }
    if (j < 0 || j > ny) {
      if (g.bc[2] == NBL_BC_PERIODIC) { j = wrapi(j, ny); continue; }
      const int jb = j < 0 ? 0 : ny;
      if (g.bc[j < 0 ? 2 : 3] == NBL_BC_OUTLET) return R.fold(v[I.v(i, jb, k)]);
      if (!R.push(v[I.v(i, jb, k)])) break;
      j = 2 * jb - j;
      continue;
    }
    return R.fold(v[I.v(i, j, k)]);
  }
  return 0.0;
}

__device__ double ghostW(const DevGrid& g, const double* __restrict__ w, int i, int j, int k) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Idx I{nx, ny, nz};
  Refl R;
  for (int guard = 0; guard < 12; ++guard) {
    if (j < 0 || j >= ny) {
      if (g.bc[2] == NBL_BC_PERIODIC) { j = wrapi(j, ny); continue; }
      if (!R.push(g.wall[j < 0 ? 2 : 3][2])) break;
      j = j < 0 ? -1 - j : 2 * ny - 1 - j;
      continue;
    }
    if (i < 0 || i >= nx) {
      if (g.bc[0] == NBL_BC_PERIODIC) { i = wrapi(i, nx); continue; }
      const int f = i < 0 ? 0 : 1;
      if (g.bc[f] == NBL_BC_OUTLET) return R.fold(w[I.w(i < 0 ? 0 : nx - 1, j, k)]);
      if (!R.push(g.wall[f][2])) break;
      i = i < 0 ? -1 - i : 2 * nx - 1 - i;
      continue;
    }
    if (k < 0 || k > nz) {
      if (g.bc[4] == NBL_BC_PERIODIC) { k = wrapi(k, nz); continue; }
      const int kb = k < 0 ? 0 : nz;
      if (!R.push(w[I.w(i, j, kb)])) break;
      k = 2 * kb - k;
      continue;
    }
    return R.fold(w[I.w(i, j, k)]);
  }
  return 0.0;
}


__device__ __forceinline__ double recon(int sc, double adv, double qm2, double qm1, double q0,
                                        double qp1, double qp2, double qp3) {
  if (sc == NBL_CONV_CENTRAL) return 0.5 * (q0 + qp1);
  return adv >= 0.0 ? weno5(qm2, qm1, q0, qp1, qp2) : weno5(qp3, qp2, qp1, q0, qm1);
}

// Stencil reader over the padded copies: element (i,j,k) of a component,
// ghost layers included, at p[i + sx*j + sxy*k] (p points at element (0,0,0)).
struct PadRd {
  const double *u, *v, *w;
  int ux, uxy, vx, vxy, wx, wxy;
  __device__ __forceinline__ double U(int i, int j, int k) const { return u[i + ux * j + uxy * k]; }
  __device__ __forceinline__ double V(int i, int j, int k) const { return v[i + vx * j + vxy * k]; }
  __device__ __forceinline__ double W(int i, int j, int k) const { return w[i + wx * j + wxy * k]; }
};

// u-momentum at x-face (i,j,k) (ns_solver.cpp:323-372).
template <int sc, class Rd>
__device__ __forceinline__ double rhsU(const NsArgs& A, const Rd& r, int i, int j, int k) {
  const DevGrid& g = A.g;
  const bool flat = g.nz == 1;
  const double aL = 0.5 * (r.U(i - 1, j, k) + r.U(i, j, k));
  const double aR = 0.5 * (r.U(i, j, k) + r.U(i + 1, j, k));
  const double qL = recon(sc, aL, r.U(i - 3, j, k), r.U(i - 2, j, k), r.U(i - 1, j, k), r.U(i, j, k),
                          r.U(i + 1, j, k), r.U(i + 2, j, k));
  const double qR = recon(sc, aR, r.U(i - 2, j, k), r.U(i - 1, j, k), r.U(i, j, k), r.U(i + 1, j, k),
                          r.U(i + 2, j, k), r.U(i + 3, j, k));
  const double fx = (aR * qR - aL * qL) / g.dx;
  const double vT = 0.5 * (r.V(i - 1, j + 1, k) + r.V(i, j + 1, k));
  const double vB = 0.5 * (r.V(i - 1, j, k) + r.V(i, j, k));
  const double qT = recon(sc, vT, r.U(i, j - 2, k), r.U(i, j - 1, k), r.U(i, j, k), r.U(i, j + 1, k),
                          r.U(i, j + 2, k), r.U(i, j + 3, k));
  const double qB = recon(sc, vB, r.U(i, j - 3, k), r.U(i, j - 2, k), r.U(i, j - 1, k), r.U(i, j, k),
                          r.U(i, j + 1, k), r.U(i, j + 2, k));
  const double fy = (vT * qT - vB * qB) / g.dy;
  double fz = 0.0;
  if (!flat) {
    const double wT = 0.5 * (r.W(i - 1, j, k + 1) + r.W(i, j, k + 1));
    const double wB = 0.5 * (r.W(i - 1, j, k) + r.W(i, j, k));
    const double zT = recon(sc, wT, r.U(i, j, k - 2), r.U(i, j, k - 1), r.U(i, j, k), r.U(i, j, k + 1),
                            r.U(i, j, k + 2), r.U(i, j, k + 3));
    const double zB = recon(sc, wB, r.U(i, j, k - 3), r.U(i, j, k - 2), r.U(i, j, k - 1), r.U(i, j, k),
                            r.U(i, j, k + 1), r.U(i, j, k + 2));
    fz = (wT * zT - wB * zB) / g.dz;
  }
  double lap = (r.U(i + 1, j, k) - 2 * r.U(i, j, k) + r.U(i - 1, j, k)) / (g.dx * g.dx) +
               (r.U(i, j + 1, k) - 2 * r.U(i, j, k) + r.U(i, j - 1, k)) / (g.dy * g.dy);
  if (!flat) lap += (r.U(i, j, k + 1) - 2 * r.U(i, j, k) + r.U(i, j, k - 1)) / (g.dz * g.dz);
  return -(fx + fy + fz) + A.nu * lap + A.bodyForce[0];
}

// v-momentum at y-face (i,j,k) (ns_solver.cpp:377-424).
template <int sc, class Rd>
__device__ __forceinline__ double rhsV(const NsArgs& A, const Rd& r, int i, int j, int k) {
  const DevGrid& g = A.g;
  const bool flat = g.nz == 1;
  const double bB = 0.5 * (r.V(i, j - 1, k) + r.V(i, j, k));
  const double bT = 0.5 * (r.V(i, j, k) + r.V(i, j + 1, k));
  const double qB = recon(sc, bB, r.V(i, j - 3, k), r.V(i, j - 2, k), r.V(i, j - 1, k), r.V(i, j, k),
                          r.V(i, j + 1, k), r.V(i, j + 2, k));
  const double qT = recon(sc, bT, r.V(i, j - 2, k), r.V(i, j - 1, k), r.V(i, j, k), r.V(i, j + 1, k),
                          r.V(i, j + 2, k), r.V(i, j + 3, k));
  const double fy = (bT * qT - bB * qB) / g.dy;
  const double uR = 0.5 * (r.U(i + 1, j - 1, k) + r.U(i + 1, j, k));
  const double uL = 0.5 * (r.U(i, j - 1, k) + r.U(i, j, k));
  const double qR = recon(sc, uR, r.V(i - 2, j, k), r.V(i - 1, j, k), r.V(i, j, k), r.V(i + 1, j, k),
                          r.V(i + 2, j, k), r.V(i + 3, j, k));
  const double qL = recon(sc, uL, r.V(i - 3, j, k), r.V(i - 2, j, k), r.V(i - 1, j, k), r.V(i, j, k),
                          r.V(i + 1, j, k), r.V(i + 2, j, k));
  const double fx = (uR * qR - uL * qL) / g.dx;
  double fz = 0.0;
  if (!flat) {
    const double wR = 0.5 * (r.W(i, j - 1, k + 1) + r.W(i, j, k + 1));
    const double wL = 0.5 * (r.W(i, j - 1, k) + r.W(i, j, k));
    const double zR = recon(sc, wR, r.V(i, j, k - 2), r.V(i, j, k - 1), r.V(i, j, k), r.V(i, j, k + 1),
                            r.V(i, j, k + 2), r.V(i, j, k + 3));
    const double zL = recon(sc, wL, r.V(i, j, k - 3), r.V(i, j, k - 2), r.V(i, j, k - 1), r.V(i, j, k),
                            r.V(i, j, k + 1), r.V(i, j, k + 2));
    fz = (wR * zR - wL * zL) / g.dz;
  }
  double lap = (r.V(i + 1, j, k) - 2 * r.V(i, j, k) + r.V(i - 1, j, k)) / (g.dx * g.dx) +
               (r.V(i, j + 1, k) - 2 * r.V(i, j, k) + r.V(i, j - 1, k)) / (g.dy * g.dy);
  if (!flat) lap += (r.V(i, j, k + 1) - 2 * r.V(i, j, k) + r.V(i, j, k - 1)) / (g.dz * g.dz);
  return -(fx + fy + fz) + A.nu * lap + A.bodyForce[1];
}

// w-momentum at z-face (i,j,k) (ns_solver.cpp:427-472).
template <int sc, class Rd>
__device__ __forceinline__ double rhsW(const NsArgs& A, const Rd& r, int i, int j, int k) {
  const DevGrid& g = A.g;
  const double cB = 0.5 * (r.W(i, j, k - 1) + r.W(i, j, k));
  const double cT = 0.5 * (r.W(i, j, k) + r.W(i, j, k + 1));
  const double qB = recon(sc, cB, r.W(i, j, k - 3), r.W(i, j, k - 2), r.W(i, j, k - 1), r.W(i, j, k),
                          r.W(i, j, k + 1), r.W(i, j, k + 2));
  const double qT = recon(sc, cT, r.W(i, j, k - 2), r.W(i, j, k - 1), r.W(i, j, k), r.W(i, j, k + 1),
                          r.W(i, j, k + 2), r.W(i, j, k + 3));
  const double fz = (cT * qT - cB * qB) / g.dz;
  const double uR = 0.5 * (r.U(i + 1, j, k - 1) + r.U(i + 1, j, k));
  const double uL = 0.5 * (r.U(i, j, k - 1) + r.U(i, j, k));
  const double xR = recon(sc, uR, r.W(i - 2, j, k), r.W(i - 1, j, k), r.W(i, j, k), r.W(i + 1, j, k),
                          r.W(i + 2, j, k), r.W(i + 3, j, k));
  const double xL = recon(sc, uL, r.W(i - 3, j, k), r.W(i - 2, j, k), r.W(i - 1, j, k), r.W(i, j, k),
                          r.W(i + 1, j, k), r.W(i + 2, j, k));
  const double fx = (uR * xR - uL * xL) / g.dx;
  const double vR = 0.5 * (r.V(i, j + 1, k - 1) + r.V(i, j + 1, k));
  const double vL = 0.5 * (r.V(i, j, k - 1) + r.V(i, j, k));
  const double yR = recon(sc, vR, r.W(i, j - 2, k), r.W(i, j - 1, k), r.W(i, j, k), r.W(i, j + 1, k),
                          r.W(i, j + 2, k), r.W(i, j + 3, k));
  const double yL = recon(sc, vL, r.W(i, j - 3, k), r.W(i, j - 2, k), r.W(i, j - 1, k), r.W(i, j, k),
                          r.W(i, j + 1, k), r.W(i, j + 2, k));
  const double fy = (vR * yR - vL * yL) / g.dy;
  const double lap = (r.W(i + 1, j, k) - 2 * r.W(i, j, k) + r.W(i - 1, j, k)) / (g.dx * g.dx) +
                     (r.W(i, j + 1, k) - 2 * r.W(i, j, k) + r.W(i, j - 1, k)) / (g.dy * g.dy) +
                     (r.W(i, j, k + 1) - 2 * r.W(i, j, k) + r.W(i, j, k - 1)) / (g.dz * g.dz);
  return -(fx + fy + fz) + A.nu * lap + A.bodyForce[2];
}

// ---------------------------------------------------------------------------
// Padded copies.  Each component array gets kPadHalo ghost layers on every
// side, filled with what the reference's ghost accessors return for those
// indices (ns_solver.cpp:116-219); the ghost value of an index depends only on
// the index and the stored array, so every stencil that reads it sees the same
// value it would in the reference.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void compDims(const DevGrid& g, int c, int& X, int& Y, int& Z) {
  X = g.nx + (c == 0);
  Y = g.ny + (c == 1);
  Z = g.nz + (c == 2);
}

// Interior copy: block z indexes the planes of u, then v, then w.
__global__ void __launch_bounds__(kBlock) k_pad_copy(NsArgs A, const double* __restrict__ u,
                                                     const double* __restrict__ v, const double* __restrict__ w,
                                                     PadBufs pb) {
  const DevGrid& g = A.g;
  int c = 0, k = blockIdx.z;
  if (k >= g.nz) {
    k -= g.nz;
    c = 1;
    if (k >= g.nz) {
      k -= g.nz;
      c = 2;
    }
  }
  int X, Y, Z;
  compDims(g, c, X, Y, Z);
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= X || j >= Y) return;
  const double* src = c == 0 ? u : (c == 1 ? v : w);
  double* dst = pb.p[c] + pb.origin[c];
  dst[i + pb.sx[c] * j + pb.sxy[c] * k] = src[i + X * (j + Y * k)];
}

// Ghost layers: one thread per halo element of u, v, w (k-planes, then
// j-rows of in-range planes, then i-ends of in-range rows).
__global__ void __launch_bounds__(kBlock) k_pad_halo(const DevGrid* __restrict__ gd, const double* __restrict__ u,
                                                     const double* __restrict__ v, const double* __restrict__ w,
                                                     PadBufs pb) {
  const DevGrid& g = *gd;
  unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int H = kPadHalo;
  for (int c = 0; c < 3; ++c) {
    int X, Y, Z;
    compDims(g, c, X, Y, Z);
    const unsigned PX = X + 2 * H, PY = Y + 2 * H;
    const unsigned nk = PX * PY * (2 * H), nj = PX * (2 * H) * unsigned(Z), ni = (2 * H) * unsigned(Y) * Z;
    if (t >= nk + nj + ni) {
      t -= nk + nj + ni;
      continue;
    }
    int i, j, k;
    if (t < nk) {
      i = int(t % PX) - H;
      const unsigned q = t / PX;
      j = int(q % PY) - H;
      const int kk = int(q / PY);
      k = kk < H ? kk - H : Z + (kk - H);
    } else if (t < nk + nj) {
      t -= nk;
      i = int(t % PX) - H;
      const unsigned q = t / PX;
      const int jj = int(q % (2 * H));
      j = jj < H ? jj - H : Y + (jj - H);
      k = int(q / (2 * H));
    } else {
      t -= nk + nj;
      const int ii = int(t % (2 * H));
      i = ii < H ? ii - H : X + (ii - H);
      const unsigned q = t / (2 * H);
      j = int(q % unsigned(Y));
      k = int(q / unsigned(Y));
    }
    // Stencils of active faces leave the array along one axis at a time, so
    // edge and corner halo elements are never read.  They are not resolved:
    // the ghost resolvers assume the other indices are in range (an outlet
    // copy would read outside the array otherwise).
    const int oob = (i < 0 || i >= X) + (j < 0 || j >= Y) + (k < 0 || k >= Z);
    double val = 0.0;
    if (oob == 1) {
      if (c == 0) val = ghostU(g, u, i, j, k);
      else if (c == 1) val = ghostV(g, v, i, j, k);
      else val = ghostW(g, w, i, j, k);
    }
    pb.p[c][pb.origin[c] + i + pb.sx[c] * j + pb.sxy[c] * k] = val;
    return;
  }
}

// SSP-RK3 stage combine of one face (ns_solver.cpp:776-794): stage 0 returns
// the right-hand side itself.
__device__ __forceinline__ double rkCombine(int stage, double dt, const double* __restrict__ a0, int t, double xc,
                                           double R) {
  if (stage == 0) return R;
  if (stage == 1) return a0[t] + dt * R;
  if (stage == 2) return 0.75 * a0[t] + 0.25 * (xc + dt * R);
  return (1.0 / 3.0) * a0[t] + (2.0 / 3.0) * (xc + dt * R);
}

// RHS + SSP-RK3 stage combine (ns_solver.cpp:776-794), one thread per face,
// block z indexing the planes of u, then v, then w.  Stencils read the padded
// copies; u0 and the output are the compact arrays.
template <int SCHEME>
__global__ void __launch_bounds__(kBlock) k_rhs_rk(NsArgs A, PadBufs pb, const double* __restrict__ u0,
                                                   const double* __restrict__ v0, const double* __restrict__ w0,
                                                   double* __restrict__ uo, double* __restrict__ vo,
                                                   double* __restrict__ wo, int stage,
                                                   const StepCtl* __restrict__ sc) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Blk3 B = blk3Tiled();
  int comp = 0, k = B.z;
  if (k >= nz) {
    k -= nz;
    comp = 1;
    if (k >= nz) {
      k -= nz;
      comp = 2;
    }
  }
  int X, Y, Z;
  compDims(g, comp, X, Y, Z);
  const int i = B.x * wgSizeX() + int(threadIdx.x), j = B.y * wgSizeY() + int(threadIdx.y);
  if (i >= X || j >= Y) return;
  const PadRd r{pb.p[0] + pb.origin[0], pb.p[1] + pb.origin[1], pb.p[2] + pb.origin[2],
                pb.sx[0], pb.sxy[0], pb.sx[1], pb.sxy[1], pb.sx[2], pb.sxy[2]};
  double R = 0.0, xc;
  const int t = i + X * (j + Y * k);
  if (comp == 0) {
    xc = r.U(i, j, k);
    if (i >= 1 && i < nx && !solidAt(g, i - 1, j, k) && !solidAt(g, i, j, k)) R = rhsU<SCHEME>(A, r, i, j, k);
  } else if (comp == 1) {
    xc = r.V(i, j, k);
    if (j >= 1 && j < ny && !solidAt(g, i, j - 1, k) && !solidAt(g, i, j, k)) R = rhsV<SCHEME>(A, r, i, j, k);
  } else {
    xc = r.W(i, j, k);
    if (nz > 1 && k >= 1 && k < nz && !solidAt(g, i, j, k - 1) && !solidAt(g, i, j, k)) R = rhsW<SCHEME>(A, r, i, j, k);
  }
  const double* a0 = comp == 0 ? u0 : (comp == 1 ? v0 : w0);
  double* ao = comp == 0 ? uo : (comp == 1 ? vo : wo);
  ao[t] = rkCombine(stage, ldA(&sc->dt), a0, t, xc, R);
}

// ---------------------------------------------------------------------------
// WENO-5 right-hand side in bricks.  The convective term of a face is a
// difference of two flux products per axis, adv * recon(adv, 6 values), and
// the upper product of face p along axis d is, operand for operand, the lower
// product of face p + e_d (ns_solver.cpp:323-472 evaluate both at every face).
// A brick is 64 faces along i by 4 rows along j of one component, marching
// through `planes` consecutive k: each face evaluates its lower x and y
// products into LDS and its upper z product in a register (the next plane's
// lower one); one wave adds the fifth row's y products and another the 65th
// column's x products.  That is 3.5 + 1/planes reconstructions per face where
// the per-face form evaluates 6; the values and the expression that combines
// them are the per-face form's, so results are bit-identical.
//
// The brick covers the faces the reference can make active (u: 1 <= i < nx,
// v: 1 <= j < ny, w: 1 <= k < nz); the rest, whose right-hand side is zero,
// are written by trailing blocks.  Every stencil stays inside the kPadHalo
// ghost layers with at most one index out of range, as the per-face form's.
// ---------------------------------------------------------------------------
// Lower flux product along an axis at face element pF of the advected
// component F: adv = (A[pA - ecA] + A[pA]) / 2 from the advecting component,
// reconstructed upwind from the six F values around the face.
__device__ __forceinline__ double weno5Flux(const double* __restrict__ F, int pF, int sF,
                                            const double* __restrict__ Av, int pA, int ecA) {
  const double adv = 0.5 * (Av[pA - ecA] + Av[pA]);
  const double q = adv >= 0.0 ? weno5(F[pF - 3 * sF], F[pF - 2 * sF], F[pF - sF], F[pF], F[pF + sF])
                              : weno5(F[pF + 2 * sF], F[pF + sF], F[pF], F[pF - sF], F[pF - 2 * sF]);
  return adv * q;
}

__global__ void __launch_bounds__(kBrickX * kBrickY)
    k_rhs_brick(NsArgs A, PadBufs pb, BrickArgs BA, const double* __restrict__ u0, const double* __restrict__ v0,
                const double* __restrict__ w0, double* __restrict__ uo, double* __restrict__ vo,
                double* __restrict__ wo, int stage, const StepCtl* __restrict__ sc) {
  __shared__ double lx[2][kBrickY][kBrickX + 1];
  __shared__ double ly[2][kBrickY + 1][kBrickX];
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool flat = nz == 1;
  const int b = xcdTile(int(blockIdx.x), int(gridDim.x));
  const double dt = ldA(&sc->dt);
  if (b >= BA.tileBlocks) {
    // faces with a zero right-hand side: u at i = 0, nx; v at j = 0, ny; w at k = 0, nz
    long long q = (long long)(b - BA.tileBlocks) * (kBrickX * kBrickY) + threadIdx.x + kBrickX * threadIdx.y;
    int c = 0;
    while (c < 3 && q >= BA.edge[c]) q -= BA.edge[c++];
    if (c == 3) return;
    const int s = int(q & 1);
    const long long h = q >> 1;
    int i, j, k;
    if (c == 0) { i = s ? nx : 0; j = int(h % ny); k = int(h / ny); }
    else if (c == 1) { j = s ? ny : 0; i = int(h % nx); k = int(h / nx); }
    else { k = s ? nz : 0; i = int(h % nx); j = int(h / nx); }
    const int X = nx + (c == 0), Y = ny + (c == 1);
    const int t = i + X * (j + Y * k);
    const double xc = pb.p[c][pb.origin[c] + i + pb.sx[c] * j + pb.sxy[c] * k];
    const double* a0 = c == 0 ? u0 : (c == 1 ? v0 : w0);
    double* ao = c == 0 ? uo : (c == 1 ? vo : wo);
    ao[t] = rkCombine(stage, dt, a0, t, xc, 0.0);
    return;
  }
  int c = 0, tb = b;
  while (tb >= BA.c[c].blocks) tb -= BA.c[c++].blocks;
  const BrickComp& C = BA.c[c];
  const int bi = tb % C.tx, bj = (tb / C.tx) % C.ty, bk = tb / (C.tx * C.ty);
  const int l = threadIdx.x, r = threadIdx.y;
  const int i0 = C.ox + bi * kBrickX, j0 = C.oy + bj * kBrickY, k0 = C.oz + bk * BA.planes;
  const int xe = C.ox + C.ex, ye = C.oy + C.ey, ke = min(k0 + BA.planes, C.oz + C.ez);
  const int i = i0 + l, j = j0 + r;
  const bool inX = i < xe, inY = j < ye, in = inX && inY;

  // F: the advected component; Ad: the component advecting along axis d, with
  // ec[d] its stride along this component's axis (the face-average offset)
  const double* F = pb.p[c] + pb.origin[c];
  const int fsx = pb.sx[c], fsxy = pb.sxy[c];
  const double* Ax = pb.p[0] + pb.origin[0];
  const double* Ay = pb.p[1] + pb.origin[1];
  const double* Az = pb.p[2] + pb.origin[2];
  const int ecx = c == 0 ? 1 : (c == 1 ? pb.sx[0] : pb.sxy[0]);
  const int ecy = c == 0 ? 1 : (c == 1 ? pb.sx[1] : pb.sxy[1]);
  const int ecz = c == 0 ? 1 : (c == 1 ? pb.sx[2] : pb.sxy[2]);
  auto fI = [&](int ii, int jj, int kk) { return ii + fsx * jj + fsxy * kk; };
  auto xI = [&](int ii, int jj, int kk) { return ii + pb.sx[0] * jj + pb.sxy[0] * kk; };
  auto yI = [&](int ii, int jj, int kk) { return ii + pb.sx[1] * jj + pb.sxy[1] * kk; };
  auto zI = [&](int ii, int jj, int kk) { return ii + pb.sx[2] * jj + pb.sxy[2] * kk; };
  auto fluxX = [&](int ii, int jj, int kk) { return weno5Flux(F, fI(ii, jj, kk), 1, Ax, xI(ii, jj, kk), ecx); };
  auto fluxY = [&](int ii, int jj, int kk) { return weno5Flux(F, fI(ii, jj, kk), fsx, Ay, yI(ii, jj, kk), ecy); };
  auto fluxZ = [&](int ii, int jj, int kk) { return weno5Flux(F, fI(ii, jj, kk), fsxy, Az, zI(ii, jj, kk), ecz); };

  const int X = nx + (c == 0), Y = ny + (c == 1);
  const double* a0 = c == 0 ? u0 : (c == 1 ? v0 : w0);
  double* ao = c == 0 ? uo : (c == 1 ? vo : wo);
  const double idx2 = g.dx * g.dx, idy2 = g.dy * g.dy, idz2 = g.dz * g.dz;
  double pzLo = 0.0;
  if (!flat && in) pzLo = fluxZ(i, j, k0);
  for (int k = k0; k < ke; ++k) {
    const int buf = (k - k0) & 1;
    if (i <= xe && inY) lx[buf][r][l] = fluxX(i, j, k);
    if (inX && j <= ye) ly[buf][r][l] = fluxY(i, j, k);
    if (r == 0 && inX && j0 + kBrickY <= ye) ly[buf][kBrickY][l] = fluxY(i, j0 + kBrickY, k);
    if (r == 1 && l < kBrickY && i0 + kBrickX <= xe && j0 + l < ye)
      lx[buf][l][kBrickX] = fluxX(i0 + kBrickX, j0 + l, k);
    double pzHi = 0.0;
    if (!flat && in) pzHi = fluxZ(i, j, k + 1);
    __syncthreads();
    if (in) {
      const int p = fI(i, j, k);
      const double xc = F[p];
      double R = 0.0;
      const bool act = !solidAt(g, i - (c == 0), j - (c == 1), k - (c == 2)) && !solidAt(g, i, j, k);
      if (act) {
        const double fx = (lx[buf][r][l + 1] - lx[buf][r][l]) / g.dx;
        const double fy = (ly[buf][r + 1][l] - ly[buf][r][l]) / g.dy;
        const double fz = flat ? 0.0 : (pzHi - pzLo) / g.dz;
        double lap = (F[p + 1] - 2 * xc + F[p - 1]) / idx2 + (F[p + fsx] - 2 * xc + F[p - fsx]) / idy2;
        if (!flat) lap += (F[p + fsxy] - 2 * xc + F[p - fsxy]) / idz2;
        R = -(fx + fy + fz) + A.nu * lap + A.bodyForce[c];
      }
      const int t = i + X * (j + Y * k);
      ao[t] = rkCombine(stage, dt, a0, t, xc, R);
    }
    pzLo = pzHi;
  }
}

// Boundary planes of applyVelocityBC (ns_solver.cpp:224-276).
__global__ void k_bc_planes(NsArgs A, double* __restrict__ u, double* __restrict__ v,
                            double* __restrict__ w, int outflow) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Idx I{nx, ny, nz};
  const long long px = 2LL * ny * nz, py = 2LL * nx * nz, pz = nz > 1 ? 2LL * nx * ny : 0;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < px) {
    const int s = int(t & 1);
    const long long q = t >> 1;
    const int j = int(q % ny), k = int(q / ny);
    const int i = s == 0 ? 0 : nx;
    const int ty = g.bc[s];
    if (ty == NBL_BC_WALL) u[I.u(i, j, k)] = 0.0;
    else if (ty == NBL_BC_INLET) u[I.u(i, j, k)] = inletComp(g, s, j, 0);
    else if (ty == NBL_BC_OUTLET && outflow) u[I.u(i, j, k)] = u[I.u(s == 0 ? 1 : nx - 1, j, k)];
    return;
  }
  t -= px;
  if (t < py) {
    const int s = int(t & 1);
    const long long q = t >> 1;
    const int i = int(q % nx), k = int(q / nx);
    const int f = 2 + s, j = s == 0 ? 0 : ny;
    const int ty = g.bc[f];
    if (ty == NBL_BC_WALL) v[I.v(i, j, k)] = 0.0;
    else if (ty == NBL_BC_INLET) v[I.v(i, j, k)] = inletComp(g, f, i, 1);
    else if (ty == NBL_BC_OUTLET && outflow) v[I.v(i, j, k)] = v[I.v(i, s == 0 ? 1 : ny - 1, k)];
    return;
  }
  t -= py;
  if (t < pz) {
    const int s = int(t & 1);
    const long long q = t >> 1;
    const int i = int(q % nx), j = int(q / nx);
    const int f = 4 + s, k = s == 0 ? 0 : nz;
    const int ty = g.bc[f];
    if (ty == NBL_BC_WALL) w[I.w(i, j, k)] = 0.0;
    else if (ty == NBL_BC_INLET) w[I.w(i, j, k)] = inletComp(g, f, i, 2);
    else if (ty == NBL_BC_OUTLET && outflow) w[I.w(i, j, k)] = w[I.w(i, j, s == 0 ? 1 : nz - 1)];
  }
}

// Solid-touching faces (precomputed lists) are zeroed (ns_solver.cpp:277-308).
__global__ void k_bc_solid(double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                           const int* __restrict__ list, int nu, int nv, int nw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nu) u[list[t]] = 0.0;
  else if (t < nu + nv) v[list[t]] = 0.0;
  else if (t < nu + nv + nw) w[list[t]] = 0.0;
}

// b = -div u on fluid cells, 0 in solids (ns_solver.cpp:482-498).
__global__ void __launch_bounds__(kBlock) k_divergence(NsArgs A, const double* __restrict__ u,
                                                       const double* __restrict__ v,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ b) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned n = unsigned(nx) * ny * nz;
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int i = int(c % unsigned(nx));
  const unsigned q = c / unsigned(nx);
  const int j = int(q % unsigned(ny)), k = int(q / unsigned(ny));
  if (g.hasSolid && g.solid[c]) {
    b[c] = 0.0;
    return;
  }
  const Idx I{nx, ny, nz};
  double d = (u[I.u(i + 1, j, k)] - u[I.u(i, j, k)]) / g.dx + (v[I.v(i, j + 1, k)] - v[I.v(i, j, k)]) / g.dy;
  if (nz > 1) d += (w[I.w(i, j, k + 1)] - w[I.w(i, j, k)]) / g.dz;
  b[c] = -d;
}

// phi with boundary ghosts (ns_solver.cpp:508-526).
__device__ __forceinline__ double phiAt(const DevGrid& g, const double* __restrict__ phi, int i, int j, int k) {
  int q[3] = {i, j, k};
  const int n[3] = {g.nx, g.ny, g.nz};
  double s = 1.0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    if (q[ax] >= 0 && q[ax] < n[ax]) continue;
    if (g.bc[2 * ax] == NBL_BC_PERIODIC) {
      q[ax] = wrapi(q[ax], n[ax]);
      continue;
    }
    const int f = 2 * ax + (q[ax] < 0 ? 0 : 1);
    q[ax] = q[ax] < 0 ? 0 : n[ax] - 1;
    s *= g.bc[f] == NBL_BC_OUTLET ? -1.0 : 1.0;
  }
  return s * phi[q[0] + g.nx * (q[1] + g.ny * q[2])];
}

// Face-flux correction u -= grad(phi) on non-solid faces, then the
// post-projection boundary values (applyVelocityBC without the outlet copy),
// and p = phi/dt (ns_solver.cpp:531-578).
__global__ void __launch_bounds__(kBlock) k_correct(NsArgs A, double* __restrict__ u, double* __restrict__ v,
                                                    double* __restrict__ w, const double* __restrict__ phi,
                                                    double* __restrict__ p, const StepCtl* __restrict__ sc) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned nu = unsigned(nx + 1) * ny * nz, nv = unsigned(nx) * (ny + 1) * nz;
  const unsigned nw = nz > 1 ? unsigned(nx) * ny * (nz + 1) : 0;
  const unsigned nc = unsigned(nx) * ny * nz;
  unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nu) {
    const int i = int(t % unsigned(nx + 1));
    const unsigned q = t / unsigned(nx + 1);
    const int j = int(q % unsigned(ny)), k = int(q / unsigned(ny));
    if (solidAt(g, i - 1, j, k) || solidAt(g, i, j, k)) {
      u[t] = 0.0;
      return;
    }
    double val = u[t] - (phiAt(g, phi, i, j, k) - phiAt(g, phi, i - 1, j, k)) / g.dx;
    if (i == 0 || i == nx) {
      const int f = i == 0 ? 0 : 1;
      if (g.bc[f] == NBL_BC_WALL) val = 0.0;
      else if (g.bc[f] == NBL_BC_INLET) val = inletComp(g, f, j, 0);
    }
    u[t] = val;
    return;
  }
  t -= nu;
  if (t < nv) {
    const int i = int(t % unsigned(nx));
    const unsigned q = t / unsigned(nx);
    const int j = int(q % unsigned(ny + 1)), k = int(q / unsigned(ny + 1));
    if (solidAt(g, i, j - 1, k) || solidAt(g, i, j, k)) {
      v[t] = 0.0;
      return;
    }
    double val = v[t] - (phiAt(g, phi, i, j, k) - phiAt(g, phi, i, j - 1, k)) / g.dy;
    if (j == 0 || j == ny) {
      const int f = j == 0 ? 2 : 3;
      if (g.bc[f] == NBL_BC_WALL) val = 0.0;
      else if (g.bc[f] == NBL_BC_INLET) val = inletComp(g, f, i, 1);
    }
    v[t] = val;
    return;
  }
  t -= nv;
  if (t < nw) {
    const int i = int(t % unsigned(nx));
    const unsigned q = t / unsigned(nx);
    const int j = int(q % unsigned(ny)), k = int(q / unsigned(ny));
    if (solidAt(g, i, j, k - 1) || solidAt(g, i, j, k)) {
      w[t] = 0.0;
      return;
    }
    double val = w[t] - (phiAt(g, phi, i, j, k) - phiAt(g, phi, i, j, k - 1)) / g.dz;
    if (k == 0 || k == nz) {
      const int f = k == 0 ? 4 : 5;
      if (g.bc[f] == NBL_BC_WALL) val = 0.0;
      else if (g.bc[f] == NBL_BC_INLET) val = inletComp(g, f, i, 2);
    }
    w[t] = val;
    return;
  }
  t -= nw;
  if (t < nc) p[t] = phi[t] / ldA(&sc->dt);
}

// CFL time step (ns_solver.cpp:581-603): per-block max over fluid cells, then
// k_cfl_fin folds the partials and sets dt.
constexpr int kNsBlocks = 4096;  // cap on blocks of the NS reductions (grid-stride)

__device__ __forceinline__ double viscTerm(const NsArgs& A) {
  const DevGrid& g = A.g;
  const bool flat = g.nz == 1;
  return 2.0 * A.nu * (1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy) + (flat ? 0.0 : 1.0 / (g.dz * g.dz)));
}

__global__ void __launch_bounds__(kBlock) k_cfl(NsArgs A, const double* __restrict__ u,
                                                const double* __restrict__ v, const double* __restrict__ w,
                                                RedBuf rb) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool flat = nz == 1;
  const unsigned n = unsigned(nx) * ny * nz;
  const Idx I{nx, ny, nz};
  const double visc = viscTerm(A);
  double m[1] = {0.0};
  for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    if (g.hasSolid && g.solid[c]) continue;
    const int i = int(c % unsigned(nx));
    const unsigned q = c / unsigned(nx);
    const int j = int(q % unsigned(ny)), k = int(q / unsigned(ny));
    const double a = fabs(0.5 * (u[I.u(i, j, k)] + u[I.u(i + 1, j, k)]));
    const double b = fabs(0.5 * (v[I.v(i, j, k)] + v[I.v(i, j + 1, k)]));
    const double cc = flat ? 0.0 : fabs(0.5 * (w[I.w(i, j, k)] + w[I.w(i, j, k + 1)]));
    const double sv = a / g.dx + b / g.dy + (flat ? 0.0 : cc / g.dz) + visc;
    m[0] = fmax(m[0], sv);
  }
  blockToPartials<1, 1u>(m, rb.partials, rb.stride, blockIdx.x);
}

__global__ void __launch_bounds__(kFinThreads) k_cfl_fin(NsArgs A, RedBuf rb, int nparts, StepCtl* sc) {
  double t[1];
  foldPartials<1, 1u>(rb.partials, rb.stride, nparts, t);
  if (threadIdx.x != 0) return;
  const double mx = fmax(viscTerm(A) + 1e-300, t[0]);
  stA(&sc->cfl_max, mx);
  stA(&sc->dt, A.cfl / mx);
}

__global__ void k_set_dt(StepCtl* sc, double dt) { stA(&sc->dt, dt); }

__global__ void k_set_step(StepCtl* sc, int step, double t, double dt) {
  stA(&sc->step, step);
  stA(&sc->t, t);
  stA(&sc->dt, dt);
  stA(&sc->pad_, 0);
  stA(&sc->cfl_max, 0.0);
}

__global__ void k_fill(double* a, long long n, double v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// End-of-step diagnostics (ns_solver.cpp:797-823, 605-675, 703-763).
__global__ void __launch_bounds__(kBlock) k_diagnostics(NsArgs A, const double* __restrict__ u,
                                                        const double* __restrict__ v,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ u0,
                                                        const double* __restrict__ v0,
                                                        const double* __restrict__ w0,
                                                        const double* __restrict__ p, StepCtl* sc,
                                                        DevReport* out, RedBuf rb) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Idx I{nx, ny, nz};
  const unsigned nu = unsigned(nx + 1) * ny * nz, nv = unsigned(nx) * (ny + 1) * nz;
  const unsigned nw = unsigned(nx) * ny * (nz + 1), nc = unsigned(nx) * ny * nz;
  const unsigned ntot = nu + nv + nw + nc;
  const double ax = g.dy * g.dz, ay = g.dx * g.dz, az = g.dx * g.dy;
  const double mu = A.nu * A.density;
  // acc: 0 du^2, 1 div^2, 2 fluid count, 3 mass net, 4 fx, 5 fy, 6 max|div|
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < ntot; t += gridDim.x * blockDim.x) {
    if (t < nu) {
      const double d = u[t] - u0[t];
      acc[0] += d * d;
      const int i = int(t % unsigned(nx + 1));
      if (i == 0) acc[3] += (u[t] - u[t + nx]) * ax;
      continue;
    }
    unsigned q = t - nu;
    if (q < nv) {
      const double d = v[q] - v0[q];
      acc[0] += d * d;
      const int j = int((q / unsigned(nx)) % unsigned(ny + 1));
      if (j == 0) acc[3] += (v[q] - v[q + unsigned(nx) * ny]) * ay;
      continue;
    }
    q -= nv;
    if (q < nw) {
      const double d = w[q] - w0[q];
      acc[0] += d * d;
      if (nz > 1 && q < unsigned(nx) * ny) acc[3] += (w[q] - w[q + unsigned(nx) * ny * nz]) * az;
      continue;
    }
    q -= nw;
    const int i = int(q % unsigned(nx));
    const unsigned r = q / unsigned(nx);
    const int j = int(r % unsigned(ny)), k = int(r / unsigned(ny));
    if (solidAt(g, i, j, k)) continue;
    double dv = (u[I.u(i + 1, j, k)] - u[I.u(i, j, k)]) / g.dx + (v[I.v(i, j + 1, k)] - v[I.v(i, j, k)]) / g.dy;
    if (nz > 1) dv += (w[I.w(i, j, k + 1)] - w[I.w(i, j, k)]) / g.dz;
    acc[1] += dv * dv;
    acc[2] += 1.0;
    acc[6] = fmax(acc[6], fabs(dv));
    if (g.hasSolid) {
      const double pf = p[q];
      const double uc = 0.5 * (u[I.u(i, j, k)] + u[I.u(i + 1, j, k)]);
      const double vc = 0.5 * (v[I.v(i, j, k)] + v[I.v(i, j + 1, k)]);
      const int nd = nz > 1 ? 6 : 4;
      for (int d = 0; d < nd; ++d) {
        const int si = i + (d == 0 ? -1 : d == 1 ? 1 : 0), sj = j + (d == 2 ? -1 : d == 3 ? 1 : 0);
        const int sk = k + (d == 4 ? -1 : d == 5 ? 1 : 0);
        if (!solidAt(g, si, sj, sk)) continue;
        const double nxo = d == 0 ? 1.0 : d == 1 ? -1.0 : 0.0;
        const double nyo = d == 2 ? 1.0 : d == 3 ? -1.0 : 0.0;
        const int axis = d / 2;
        const double area = axis == 0 ? ax : axis == 1 ? ay : az;
        const double dn = axis == 0 ? g.dx : axis == 1 ? g.dy : g.dz;
        acc[4] += -pf * nxo * area;
        acc[5] += -pf * nyo * area;
        if (axis != 0) acc[4] += mu * uc / (0.5 * dn) * area;
        if (axis != 1) acc[5] += mu * vc / (0.5 * dn) * area;
      }
    }
  }
  blockToPartials<7, 1u << 6>(acc, rb.partials, rb.stride, blockIdx.x);
}

// Step report from the folded diagnostics (ns_solver.cpp:797-823).
__global__ void __launch_bounds__(kFinThreads) k_diag_fin(NsArgs A, RedBuf rb, int nparts, StepCtl* sc,
                                                          DevReport* out) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const double ay = g.dx * g.dz;
  const double nu = double(nx + 1) * ny * nz, nv = double(nx) * (ny + 1) * nz, nw = double(nx) * ny * (nz + 1);
  double tot[7];
  foldPartials<7, 1u << 6>(rb.partials, rb.stride, nparts, tot);
  if (threadIdx.x != 0) return;
  const double dt = ldA(&sc->dt);
  const double n = nu + nv + nw;
  const double mom = sqrt(tot[0] / n) / (dt * A.velocityScale + 1e-300);
  const double divl2 = tot[2] > 0.0 ? sqrt(tot[1] / tot[2]) : 0.0;
  const double mass = fabs(tot[3]) / (A.velocityScale * (ny * ay) + 1e-300);
  double cd = 0.0, cl = 0.0;
  if (g.hasSolid) {
    double aref = A.refArea;
    if (aref <= 0.0) aref = A.refLength * A.refLength;
    const double qd = 0.5 * A.density * A.velocityScale * A.velocityScale * aref + 1e-300;
    cd = tot[4] / qd;
    cl = tot[5] / qd;
  }
  const int step = ldA(&sc->step) + 1;
  const double tnow = ldA(&sc->t) + dt;
  stA(&sc->step, step);
  stA(&sc->t, tnow);
  out->step = step;
  out->t = tnow;
  out->dt = dt;
  out->cfl = A.cfl;
  out->mom = mom;
  out->divl2 = divl2;
  out->divmax = tot[6];
  out->mass = mass;
  out->cd = cd;
  out->cl = cl;
  out->cells = (unsigned long long)nx * ny * nz;
  out->accepted = (divl2 < A.tolContinuity && mass < A.tolMass) ? 1 : 0;
}

// Cell-centred velocity (ns_solver.cpp:677-701).
__global__ void k_cell_velocity(NsArgs A, const double* __restrict__ u, const double* __restrict__ v,
                                const double* __restrict__ w, int comp, double* __restrict__ out) {
  const DevGrid& g = A.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned n = unsigned(nx) * ny * nz;
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const Idx I{nx, ny, nz};
  const int i = int(c % unsigned(nx));
  const unsigned q = c / unsigned(nx);
  const int j = int(q % unsigned(ny)), k = int(q / unsigned(ny));
  double r;
  if (comp == 0) r = 0.5 * (u[I.u(i, j, k)] + u[I.u(i + 1, j, k)]);
  else if (comp == 1) r = 0.5 * (v[I.v(i, j, k)] + v[I.v(i, j + 1, k)]);
  else r = nz == 1 ? 0.0 : 0.5 * (w[I.w(i, j, k)] + w[I.w(i, j, k + 1)]);
  out[c] = r;
}

}  // namespace
}  // namespace nbl

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace nbl {

static long long faces(const DevGrid& g) {
  return (long long)(g.nx + 1) * g.ny * g.nz + (long long)g.nx * (g.ny + 1) * g.nz +
         (long long)g.nx * g.ny * (g.nz + 1);
}

void launch_bc_planes(const NsArgs& a, double* u, double* v, double* w, bool outflow, hipStream_t s) {
  const DevGrid& g = a.g;
  const long long n = 2LL * g.ny * g.nz + 2LL * g.nx * g.nz + (g.nz > 1 ? 2LL * g.nx * g.ny : 0);
  k_bc_planes<<<gridFor(n), kBlock, 0, s>>>(a, u, v, w, outflow ? 1 : 0);
  NBL_LAUNCH_CHECK();
}

void launch_bc_solid(const NsArgs& a, double* u, double* v, double* w, const int* list,
                     const int counts[3], hipStream_t s) {
  (void)a;
  const int n = counts[0] + counts[1] + counts[2];
  if (n == 0) return;
  k_bc_solid<<<gridFor(n), kBlock, 0, s>>>(u, v, w, list, counts[0], counts[1], counts[2]);
  NBL_LAUNCH_CHECK();
}

PadBufs padLayout(const DevGrid& g) {
  PadBufs pb{};
  const int H = kPadHalo;
  for (int c = 0; c < 3; ++c) {
    const long long X = g.nx + (c == 0), Y = g.ny + (c == 1), Z = g.nz + (c == 2);
    const long long PX = X + 2 * H, PY = Y + 2 * H, PZ = Z + 2 * H;
    pb.sx[c] = int(PX);
    pb.sxy[c] = int(PX * PY);
    pb.origin[c] = H + PX * (H + PY * H);
    pb.n[c] = PX * PY * PZ;
  }
  return pb;
}

static long long haloCount(const DevGrid& g) {
  const long long H = kPadHalo;
  long long n = 0;
  for (int c = 0; c < 3; ++c) {
    const long long X = g.nx + (c == 0), Y = g.ny + (c == 1), Z = g.nz + (c == 2);
    n += (X + 2 * H) * (Y + 2 * H) * 2 * H + (X + 2 * H) * 2 * H * Z + 2 * H * Y * Z;
  }
  return n;
}

void launch_rhs_rk(const NsArgs& a, const DevGrid* gd, const PadBufs& pb, const double* u0, const double* v0,
                   const double* w0, const double* u, const double* v, const double* w, double* uo, double* vo,
                   double* wo, int stage, const StepCtl* sc, hipStream_t s, int planes) {
  const DevGrid& g = a.g;
  for (int c = 0; c < 3; ++c)
    if (pb.n[c] >= (1LL << 31)) throw std::invalid_argument("grid too large for 32-bit face indexing");
  // block z walks the planes of u (nz), v (nz) and w (nz + 1)
  const Geo3 geo = geo3(g.nx + 1, g.ny + 1, 3 * g.nz + 1);
  const dim3 grid = geo.grid;
  k_pad_copy<<<grid, geo.block, 0, s>>>(a, u, v, w, pb);
  k_pad_halo<<<gridFor(haloCount(g)), kBlock, 0, s>>>(gd, u, v, w, pb);
  if (a.convection == NBL_CONV_CENTRAL) {
    k_rhs_rk<NBL_CONV_CENTRAL><<<grid, geo.block, 0, s>>>(a, pb, u0, v0, w0, uo, vo, wo, stage, sc);
  } else {
    const BrickArgs ba = brickArgs(g, planes);
    const long long edges = ba.edge[0] + ba.edge[1] + ba.edge[2];
    const long long blocks = ba.tileBlocks + (edges + kBrickX * kBrickY - 1) / (kBrickX * kBrickY);
    if (blocks >= (1LL << 31)) throw std::invalid_argument("grid too large for the right-hand side bricks");
    k_rhs_brick<<<dim3(unsigned(blocks)), dim3(kBrickX, kBrickY), 0, s>>>(a, pb, ba, u0, v0, w0, uo, vo, wo,
                                                                         stage, sc);
  }
  NBL_LAUNCH_CHECK();
}

// Brick decomposition of the WENO-5 right-hand side: per component the box of
// faces that can be active, tiled 64 x 4 x planes.  `planes` <= 0 picks the
// longest march of 4, 2 or 1 planes that still gives >= 4096 bricks (16 per
// CU): on the 192 x 64 x 64 channel, bricks of 2 planes (4608) ran the kernel
// in 90 us, of 4 in 93 us, of 8 in 117 us and of 16 in 150 us, so enough
// resident waves to hide the stencil loads count for more than the z fluxes
// a longer march saves.
BrickArgs brickArgs(const DevGrid& g, int planes) {
  if (planes > 16) throw std::invalid_argument("rhs_planes must be 0 (automatic) or 1..16");
  BrickArgs a{};
  auto tile = [&](int p) {
    a.planes = p;
    a.tileBlocks = 0;
    for (int c = 0; c < 3; ++c) {
      BrickComp& C = a.c[c];
      C.ox = c == 0;
      C.oy = c == 1;
      C.oz = c == 2;
      C.ex = g.nx - (c == 0);
      C.ey = g.ny - (c == 1);
      C.ez = g.nz - (c == 2);
      C.tx = std::max(1, (C.ex + kBrickX - 1) / kBrickX);
      C.ty = std::max(1, (C.ey + kBrickY - 1) / kBrickY);
      C.tz = std::max(1, (C.ez + p - 1) / p);
      C.blocks = (C.ex > 0 && C.ey > 0 && C.ez > 0) ? C.tx * C.ty * C.tz : 0;
      a.tileBlocks += C.blocks;
    }
  };
  if (planes > 0) {
    tile(planes);
  } else {
    for (int p : {4, 2, 1}) {
      tile(p);
      if (a.tileBlocks >= 4096) break;
    }
  }
  a.edge[0] = 2LL * g.ny * g.nz;
  a.edge[1] = 2LL * g.nx * g.nz;
  a.edge[2] = 2LL * g.nx * g.ny;
  return a;
}

void launch_divergence(const NsArgs& a, const double* u, const double* v, const double* w,
                       double* b, hipStream_t s) {
  const long long n = (long long)a.g.nx * a.g.ny * a.g.nz;
  k_divergence<<<gridFor(n), kBlock, 0, s>>>(a, u, v, w, b);
  NBL_LAUNCH_CHECK();
}

void launch_correct(const NsArgs& a, double* u, double* v, double* w, const double* phi,
                    double* p, const StepCtl* sc, hipStream_t s) {
  const long long n = faces(a.g) + (long long)a.g.nx * a.g.ny * a.g.nz;
  k_correct<<<gridFor(n), kBlock, 0, s>>>(a, u, v, w, phi, p, sc);
  NBL_LAUNCH_CHECK();
}

int ns_partials(const DevGrid& g) {
  (void)g;
  return kNsBlocks;
}

void launch_cfl(const NsArgs& a, const double* u, const double* v, const double* w, StepCtl* sc,
                double fixed_dt, RedBuf rb, hipStream_t s) {
  if (fixed_dt > 0.0) {
    k_set_dt<<<1, 1, 0, s>>>(sc, fixed_dt);
  } else {
    const long long n = (long long)a.g.nx * a.g.ny * a.g.nz;
    const int g = gridFor(n, kBlock, kNsBlocks);
    k_cfl<<<g, kBlock, 0, s>>>(a, u, v, w, rb);
    k_cfl_fin<<<1, kFinThreads, 0, s>>>(a, rb, g, sc);
  }
  NBL_LAUNCH_CHECK();
}

void launch_diagnostics(const NsArgs& a, const double* u, const double* v, const double* w,
                        const double* u0, const double* v0, const double* w0, const double* p,
                        StepCtl* sc, DevReport* out, RedBuf rb, hipStream_t s) {
  const long long n = faces(a.g) + (long long)a.g.nx * a.g.ny * a.g.nz;
  const int g = gridFor(n, kBlock, kNsBlocks);
  k_diagnostics<<<g, kBlock, 0, s>>>(a, u, v, w, u0, v0, w0, p, sc, out, rb);
  k_diag_fin<<<1, kFinThreads, 0, s>>>(a, rb, g, sc, out);
  NBL_LAUNCH_CHECK();
}

void launch_set_step(StepCtl* sc, int step, double t, double dt, hipStream_t s) {
  k_set_step<<<1, 1, 0, s>>>(sc, step, t, dt);
  NBL_LAUNCH_CHECK();
}

void launch_fill(double* a, long long n, double v, hipStream_t s) {
  if (n <= 0) return;
  k_fill<<<gridFor(n), kBlock, 0, s>>>(a, n, v);
  NBL_LAUNCH_CHECK();
}

void launch_cell_velocity(const NsArgs& a, const double* u, const double* v, const double* w,
                          int comp, double* out, hipStream_t s) {
  const long long n = (long long)a.g.nx * a.g.ny * a.g.nz;
  k_cell_velocity<<<gridFor(n), kBlock, 0, s>>>(a, u, v, w, comp, out);
  NBL_LAUNCH_CHECK();
}

}  // namespace nbl
