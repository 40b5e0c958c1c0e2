// oracle.cpp — TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Serial FP64 restatement of the reference hot path.  Built with
// -ffp-contract=off so every expression rounds exactly as the reference's
// (GCC, baseline x86-64, no FMA).  Each block cites the reference lines whose
// behaviour it restates; the code itself is written from that behaviour.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace orc {

using Vec = std::vector<double>;

// ---------------------------------------------------------------------------
// Grid (core/include/nabla/flow/grid.hpp:32-73)
// ---------------------------------------------------------------------------
struct Grid {
  int nx = 1, ny = 1, nz = 1;
  double org[3] = {0, 0, 0};
  double dx = 1, dy = 1, dz = 1;
  nbl_face_bc face[6]{};
  std::vector<uint8_t> solid;

  size_t cells() const { return size_t(nx) * size_t(ny) * size_t(nz); }
  size_t c(int i, int j, int k) const { return size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * size_t(k)); }
  bool solidAt(int i, int j, int k) const {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return false;
    return solid[c(i, j, k)] != 0;
  }
  bool per(int axis) const { return face[2 * axis].type == NBL_BC_PERIODIC; }
  int n(int axis) const { return axis == 0 ? nx : (axis == 1 ? ny : nz); }
  double h(int axis) const { return axis == 0 ? dx : (axis == 1 ? dy : dz); }
};

static Grid makeGrid(const nbl_grid_desc* d) {
  Grid g;
  g.nx = d->nx;
  g.ny = d->ny;
  g.nz = d->nz;
  for (int a = 0; a < 3; ++a) g.org[a] = d->origin[a];
  g.dx = d->dx;
  g.dy = d->dy;
  g.dz = d->dz;
  for (int f = 0; f < 6; ++f) g.face[f] = d->face[f];
  g.solid.assign(g.cells(), 0);
  if (d->solid) std::memcpy(g.solid.data(), d->solid, g.cells());
  return g;
}

static int wrapIdx(int a, int n) { return (a % n + n) % n; }

// ---------------------------------------------------------------------------
// Pressure operator (core/src/flow/poisson.cpp:18-134)
// ---------------------------------------------------------------------------
// Coupling of cell (i,j,k) across its `side` face on `axis`:
//   0 = none (Neumann wall/inlet or solid neighbour), 1 = fluid neighbour
//   `nb`, 2 = outlet Dirichlet.
static int faceKind(const Grid& g, int i, int j, int k, int axis, int side, size_t& nb) {
  int q[3] = {i, j, k};
  q[axis] += side;
  const int n = g.n(axis);
  if (q[axis] < 0 || q[axis] >= n) {
    const int fid = 2 * axis + (side > 0 ? 1 : 0);
    const int t = g.face[fid].type;
    if (t == NBL_BC_PERIODIC) {
      q[axis] = (q[axis] + n) % n;
    } else if (t == NBL_BC_OUTLET) {
      return 2;
    } else {
      return 0;
    }
  }
  if (g.solidAt(q[0], q[1], q[2])) return 0;
  nb = g.c(q[0], q[1], q[2]);
  return 1;
}

static double diagSum(const Grid& g, int i, int j, int k) {
  const double inv[3] = {1.0 / (g.dx * g.dx), 1.0 / (g.dy * g.dy), 1.0 / (g.dz * g.dz)};
  double d = 0.0;
  for (int axis = 0; axis < 3; ++axis) {
    if (axis == 2 && g.nz == 1) continue;
    for (int side = -1; side <= 1; side += 2) {
      size_t nb = 0;
      const int kind = faceKind(g, i, j, k, axis, side, nb);
      if (kind == 1) d += inv[axis];
      else if (kind == 2) d += 2.0 * inv[axis];
    }
  }
  return d;
}

static void buildDiag(const Grid& g, Vec& diag) {
  diag.assign(g.cells(), 1.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        if (g.solidAt(i, j, k)) continue;
        const double d = diagSum(g, i, j, k);
        diag[g.c(i, j, k)] = d > 0.0 ? d : 1.0;
      }
}

static bool anyOutlet(const Grid& g) {
  for (const auto& f : g.face)
    if (f.type == NBL_BC_OUTLET) return true;
  return false;
}

static void applyA(const Grid& g, const Vec& x, Vec& out) {
  const double inv[3] = {1.0 / (g.dx * g.dx), 1.0 / (g.dy * g.dy), 1.0 / (g.dz * g.dz)};
  out.assign(x.size(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t c = g.c(i, j, k);
        if (g.solidAt(i, j, k)) {
          out[c] = x[c];
          continue;
        }
        double dg = 0.0, off = 0.0;
        for (int axis = 0; axis < 3; ++axis) {
          if (axis == 2 && g.nz == 1) continue;
          for (int side = -1; side <= 1; side += 2) {
            size_t nb = 0;
            const int kind = faceKind(g, i, j, k, axis, side, nb);
            if (kind == 1) {
              dg += inv[axis];
              off += inv[axis] * x[nb];
            } else if (kind == 2) {
              dg += 2.0 * inv[axis];
            }
          }
        }
        out[c] = dg * x[c] - off;
      }
}

static double dotv(const Vec& a, const Vec& b) {
  double s = 0.0;
  for (size_t m = 0; m < a.size(); ++m) s += a[m] * b[m];
  return s;
}

// Jacobi-PCG (core/src/flow/poisson.cpp:136-252).
static nbl_poisson_report cgSolve(const Grid& g, const Vec& b, Vec& x, double tol, int maxIt) {
  const size_t N = g.cells();
  x.assign(N, 0.0);
  Vec diag(N, 1.0);
  std::vector<uint8_t> fl(N, 0);
  size_t nf = 0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const size_t c = g.c(i, j, k);
        if (g.solidAt(i, j, k)) continue;
        fl[c] = 1;
        ++nf;
        const double d = diagSum(g, i, j, k);
        diag[c] = d > 0.0 ? d : 1.0;
      }
  Vec rhs = b;
  const bool sing = !anyOutlet(g);
  if (sing && nf > 0) {
    double m = 0.0;
    for (size_t c = 0; c < N; ++c)
      if (fl[c]) m += rhs[c];
    m /= double(nf);
    for (size_t c = 0; c < N; ++c) rhs[c] = fl[c] ? rhs[c] - m : 0.0;
  }
  Vec r = rhs, z(N), p(N), Ap(N);
  for (size_t c = 0; c < N; ++c) z[c] = r[c] / diag[c];
  p = z;
  double rz = dotv(r, z);
  const double bn = std::sqrt(dotv(rhs, rhs)) + 1e-300;
  nbl_poisson_report rep{0, 0.0, 0};
  int it = 0;
  for (; it < maxIt; ++it) {
    applyA(g, p, Ap);
    const double pAp = dotv(p, Ap);
    if (std::abs(pAp) < 1e-300) break;
    const double al = rz / pAp;
    for (size_t c = 0; c < N; ++c) {
      x[c] += al * p[c];
      r[c] -= al * Ap[c];
    }
    if (std::sqrt(dotv(r, r)) <= tol * bn) {
      ++it;
      rep.converged = 1;
      break;
    }
    for (size_t c = 0; c < N; ++c) z[c] = r[c] / diag[c];
    const double rz2 = dotv(r, z);
    const double be = rz2 / rz;
    for (size_t c = 0; c < N; ++c) p[c] = z[c] + be * p[c];
    rz = rz2;
  }
  if (sing && nf > 0) {
    double m = 0.0;
    for (size_t c = 0; c < N; ++c)
      if (fl[c]) m += x[c];
    m /= double(nf);
    for (size_t c = 0; c < N; ++c)
      if (fl[c]) x[c] -= m;
  }
  applyA(g, x, Ap);
  double rr = 0.0;
  for (size_t c = 0; c < N; ++c) {
    const double e = rhs[c] - Ap[c];
    rr += e * e;
  }
  rep.iterations = it;
  rep.residual = std::sqrt(rr) / bn;
  rep.converged = (rep.converged || rep.residual <= 1e-6) ? 1 : 0;
  return rep;
}

// ---------------------------------------------------------------------------
// Geometric multigrid (core/src/flow/multigrid.cpp:22-355)
// ---------------------------------------------------------------------------
struct Level {
  Grid g;
  Vec diag;
  std::vector<uint8_t> fl;
  size_t nf = 0;
  int f[3] = {1, 1, 1};  // coarsening factor from the next-finer level
};

struct Hier {
  std::vector<Level> lv;
  bool sing = false;
  uint64_t sig = 0;
};

static uint64_t gridSig(const Grid& g) {  // FNV-1a over dims, BC kinds, solid cells
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  mix(uint64_t(g.nx));
  mix(uint64_t(g.ny));
  mix(uint64_t(g.nz));
  for (int f = 0; f < 6; ++f) mix(uint64_t(g.face[f].type));
  for (size_t i = 0; i < g.solid.size(); ++i)
    if (g.solid[i]) mix(i);
  return h;
}

static void levelMeta(Level& L) {
  buildDiag(L.g, L.diag);
  L.fl.assign(L.g.cells(), 0);
  L.nf = 0;
  for (int k = 0; k < L.g.nz; ++k)
    for (int j = 0; j < L.g.ny; ++j)
      for (int i = 0; i < L.g.nx; ++i)
        if (!L.g.solidAt(i, j, k)) {
          L.fl[L.g.c(i, j, k)] = 1;
          ++L.nf;
        }
}

static std::shared_ptr<Hier> buildHier(const Grid& g) {
  auto H = std::make_shared<Hier>();
  H->sig = gridSig(g);
  H->sing = !anyOutlet(g);
  Level L0;
  L0.g = g;
  levelMeta(L0);
  H->lv.push_back(std::move(L0));
  while (H->lv.size() < 8) {
    const Grid& fg = H->lv.back().g;
    const int fx = fg.nx >= 4 ? 2 : 1, fy = fg.ny >= 4 ? 2 : 1, fz = fg.nz >= 4 ? 2 : 1;
    if (fx == 1 && fy == 1 && fz == 1) break;
    Level C;
    C.f[0] = fx;
    C.f[1] = fy;
    C.f[2] = fz;
    Grid& cg = C.g;
    cg.nx = (fg.nx + fx - 1) / fx;
    cg.ny = (fg.ny + fy - 1) / fy;
    cg.nz = (fg.nz + fz - 1) / fz;
    cg.dx = fg.dx * double(fg.nx) / double(cg.nx);
    cg.dy = fg.dy * double(fg.ny) / double(cg.ny);
    cg.dz = fg.dz * double(fg.nz) / double(cg.nz);
    for (int a = 0; a < 3; ++a) cg.org[a] = fg.org[a];
    for (int f = 0; f < 6; ++f) cg.face[f] = fg.face[f];
    std::vector<uint8_t> any(cg.cells(), 0);
    const Level& F = H->lv.back();
    for (int k = 0; k < fg.nz; ++k)
      for (int j = 0; j < fg.ny; ++j)
        for (int i = 0; i < fg.nx; ++i)
          if (F.fl[fg.c(i, j, k)]) any[cg.c(i / fx, j / fy, k / fz)] = 1;
    cg.solid.assign(cg.cells(), 0);
    for (size_t q = 0; q < cg.solid.size(); ++q) cg.solid[q] = any[q] ? 0 : 1;
    levelMeta(C);
    H->lv.push_back(std::move(C));
    if (H->lv.back().g.cells() <= 2048) break;
  }
  return H;
}

static void projectMean(const Level& L, Vec& x) {
  if (L.nf == 0) return;
  double m = 0.0;
  for (size_t c = 0; c < x.size(); ++c) {
    if (L.fl[c]) m += x[c];
    else x[c] = 0.0;
  }
  m /= double(L.nf);
  for (size_t c = 0; c < x.size(); ++c)
    if (L.fl[c]) x[c] -= m;
}

static void jacobi(const Level& L, const Vec& b, Vec& x, int sweeps, bool sing) {
  Vec Ax(L.g.cells());
  for (int s = 0; s < sweeps; ++s) {
    applyA(L.g, x, Ax);
    for (size_t c = 0; c < x.size(); ++c)
      x[c] = L.fl[c] ? x[c] + 0.8 * (b[c] - Ax[c]) / L.diag[c] : 0.0;
    if (sing) projectMean(L, x);
  }
}

static void resid(const Level& L, const Vec& b, const Vec& x, Vec& r) {
  applyA(L.g, x, r);
  for (size_t c = 0; c < r.size(); ++c) r[c] = L.fl[c] ? b[c] - r[c] : 0.0;
}

static void restrictTo(const Level& F, const Level& C, const Vec& rf, Vec& rc, bool sing, bool project) {
  const size_t nc = C.g.cells();
  rc.assign(nc, 0.0);
  std::vector<int> cnt(nc, 0);
  for (int k = 0; k < F.g.nz; ++k)
    for (int j = 0; j < F.g.ny; ++j)
      for (int i = 0; i < F.g.nx; ++i) {
        const size_t f = F.g.c(i, j, k);
        if (!F.fl[f]) continue;
        const size_t q = C.g.c(i / C.f[0], j / C.f[1], k / C.f[2]);
        rc[q] += rf[f];
        ++cnt[q];
      }
  for (size_t q = 0; q < nc; ++q)
    if (cnt[q] > 0) rc[q] /= double(cnt[q]);
  if (sing && project) projectMean(C, rc);
}

static void prolongAdd(const Level& C, const Level& F, const Vec& ec, Vec& xf) {
  for (int k = 0; k < F.g.nz; ++k)
    for (int j = 0; j < F.g.ny; ++j)
      for (int i = 0; i < F.g.nx; ++i) {
        const size_t f = F.g.c(i, j, k);
        if (!F.fl[f]) continue;
        xf[f] += ec[C.g.c(i / C.f[0], j / C.f[1], k / C.f[2])];
      }
}

static void vcyc(Hier& H, size_t l, const Vec& b, Vec& x, int pre, int post) {
  Level& L = H.lv[l];
  if (l + 1 >= H.lv.size()) {
    cgSolve(L.g, b, x, 1e-6, 500);
    return;
  }
  jacobi(L, b, x, pre, H.sing);
  Vec r;
  resid(L, b, x, r);
  Vec rc;
  restrictTo(L, H.lv[l + 1], r, rc, H.sing, true);
  Vec ec(H.lv[l + 1].g.cells(), 0.0);
  vcyc(H, l + 1, rc, ec, pre, post);
  prolongAdd(H.lv[l + 1], L, ec, x);
  jacobi(L, b, x, post, H.sing);
}

struct Mg {
  double tol;
  int maxCyc;
  int pre = 2, post = 2;
  std::shared_ptr<Hier> cache;

  nbl_poisson_report solve(const Grid& g, const Vec& b, Vec& x) {
    const uint64_t sig = gridSig(g);
    if (!cache || cache->sig != sig) cache = buildHier(g);
    Hier& H = *cache;
    const size_t N = g.cells();
    x.assign(N, 0.0);
    Vec rhs = b;
    for (size_t c = 0; c < N; ++c)
      if (!H.lv[0].fl[c]) rhs[c] = 0.0;
    if (H.sing) projectMean(H.lv[0], rhs);
    const double bn = std::sqrt(dotv(rhs, rhs)) + 1e-300;
    auto precond = [&](const Vec& rr, Vec& zz) {
      zz.assign(N, 0.0);
      vcyc(H, 0, rr, zz, pre, post);
      if (H.sing) projectMean(H.lv[0], zz);
    };
    nbl_poisson_report rep{0, 0.0, 0};
    Vec r = rhs;
    if (std::sqrt(dotv(r, r)) <= tol * bn) {
      rep.iterations = 0;
      rep.residual = std::sqrt(dotv(r, r)) / bn;
      rep.converged = 1;
      return rep;
    }
    Vec z, p, Ap, rn, zn;
    precond(r, z);
    p = z;
    double rz = dotv(r, z);
    int it = 0;
    for (; it < maxCyc; ++it) {
      applyA(g, p, Ap);
      const double pAp = dotv(p, Ap);
      if (std::abs(pAp) < 1e-300) break;
      const double al = rz / pAp;
      rn.assign(N, 0.0);
      for (size_t c = 0; c < N; ++c) {
        x[c] += al * p[c];
        rn[c] = r[c] - al * Ap[c];
      }
      if (H.sing) projectMean(H.lv[0], rn);
      if (std::sqrt(dotv(rn, rn)) <= tol * bn) {
        ++it;
        r = rn;
        rep.converged = 1;
        break;
      }
      precond(rn, zn);
      double a1 = 0.0, a2 = 0.0;
      for (size_t c = 0; c < N; ++c) {
        a1 += zn[c] * rn[c];
        a2 += zn[c] * r[c];
      }
      const double be = (a1 - a2) / rz;
      for (size_t c = 0; c < N; ++c) p[c] = zn[c] + be * p[c];
      rz = a1;
      r = rn;
    }
    if (H.sing) projectMean(H.lv[0], x);
    applyA(g, x, Ap);
    double rr = 0.0;
    for (size_t c = 0; c < N; ++c) {
      const double e = rhs[c] - Ap[c];
      rr += e * e;
    }
    rep.iterations = it;
    rep.residual = std::sqrt(rr) / bn;
    rep.converged = (rep.converged || rep.residual <= tol) ? 1 : 0;
    return rep;
  }
};

// ---------------------------------------------------------------------------
// MAC Navier-Stokes (core/src/flow/ns_solver.cpp)
// ---------------------------------------------------------------------------
// Jiang-Shu WENO-5 value at the right face of v2 (ns_solver.cpp:15-31).
static double weno(double v0, double v1, double v2, double v3, double v4) {
  const double e = 1e-12;
  const double s0 = 13.0 / 12.0 * (v0 - 2 * v1 + v2) * (v0 - 2 * v1 + v2) +
                    0.25 * (v0 - 4 * v1 + 3 * v2) * (v0 - 4 * v1 + 3 * v2);
  const double s1 = 13.0 / 12.0 * (v1 - 2 * v2 + v3) * (v1 - 2 * v2 + v3) + 0.25 * (v1 - v3) * (v1 - v3);
  const double s2 = 13.0 / 12.0 * (v2 - 2 * v3 + v4) * (v2 - 2 * v3 + v4) +
                    0.25 * (3 * v2 - 4 * v3 + v4) * (3 * v2 - 4 * v3 + v4);
  const double w0 = 0.1 / ((e + s0) * (e + s0));
  const double w1 = 0.6 / ((e + s1) * (e + s1));
  const double w2 = 0.3 / ((e + s2) * (e + s2));
  const double sw = w0 + w1 + w2;
  const double q0 = (2 * v0 - 7 * v1 + 11 * v2) / 6.0;
  const double q1 = (-v1 + 5 * v2 + 2 * v3) / 6.0;
  const double q2 = (2 * v2 + 5 * v3 - v4) / 6.0;
  return (w0 * q0 + w1 * q1 + w2 * q2) / sw;
}

// Face value between q0 and qp1, upwinded by `adv` (ns_solver.cpp:35-42).
static double faceVal(int scheme, double adv, double qm2, double qm1, double q0, double qp1,
                      double qp2, double qp3) {
  if (scheme == NBL_CONV_CENTRAL) return 0.5 * (q0 + qp1);
  return adv >= 0.0 ? weno(qm2, qm1, q0, qp1, qp2) : weno(qp3, qp2, qp1, q0, qm1);
}

struct NS {
  Grid g;
  nbl_flow_cfg cfg{};
  double nu = 0;
  bool anySolid = false;
  Vec u, v, w, p;
  int stepN = 0;
  double t = 0, dtLast = 0;
  Mg mg;

  size_t iu(int i, int j, int k) const { return size_t(i) + size_t(g.nx + 1) * (size_t(j) + size_t(g.ny) * size_t(k)); }
  size_t iv(int i, int j, int k) const { return size_t(i) + size_t(g.nx) * (size_t(j) + size_t(g.ny + 1) * size_t(k)); }
  size_t iw(int i, int j, int k) const { return size_t(i) + size_t(g.nx) * (size_t(j) + size_t(g.ny) * size_t(k)); }

  // Inlet velocity on face `f` at transverse indices (a, b) (ns_solver.cpp:71-93):
  // the parabolic profile always uses `a` as the y index.
  double inletComp(int f, int a, int comp) const {
    const nbl_face_bc& fb = g.face[f];
    double mag = fb.inlet_speed;
    if (fb.profile == NBL_INLET_PARABOLIC) {
      const double H = g.ny * g.dy;
      const double yc = g.org[1] + 0.5 * H;
      const double y = g.org[1] + (a + 0.5) * g.dy;
      const double r = (y - yc) / (0.5 * H);
      mag = 1.5 * fb.inlet_speed * (1.0 - r * r);
    }
    const int ax = fb.inlet_axis == 0 ? 0 : (fb.inlet_axis == 1 ? 1 : 2);
    return ax == comp ? mag : 0.0;
  }

  // Ghost-aware reads of the staggered components (ns_solver.cpp:116-219).
  double U(const Vec& a, int i, int j, int k) const {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (k < 0 || k >= nz) {
      if (g.per(2)) return U(a, i, j, wrapIdx(k, nz));
      const int f = k < 0 ? 4 : 5;
      return 2.0 * g.face[f].wall_velocity[0] - U(a, i, j, k < 0 ? -1 - k : 2 * nz - 1 - k);
    }
    if (j < 0 || j >= ny) {
      if (g.per(1)) return U(a, i, wrapIdx(j, ny), k);
      const int f = j < 0 ? 2 : 3;
      return 2.0 * g.face[f].wall_velocity[0] - U(a, i, j < 0 ? -1 - j : 2 * ny - 1 - j, k);
    }
    if (i < 0 || i > nx) {
      if (g.per(0)) return U(a, wrapIdx(i, nx), j, k);
      const int ib = i < 0 ? 0 : nx;
      const int f = i < 0 ? 0 : 1;
      if (g.face[f].type == NBL_BC_OUTLET) return a[iu(ib, j, k)];
      return 2.0 * a[iu(ib, j, k)] - U(a, 2 * ib - i, j, k);
    }
    return a[iu(i, j, k)];
  }
  double V(const Vec& a, int i, int j, int k) const {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (k < 0 || k >= nz) {
      if (g.per(2)) return V(a, i, j, wrapIdx(k, nz));
      const int f = k < 0 ? 4 : 5;
      return 2.0 * g.face[f].wall_velocity[1] - V(a, i, j, k < 0 ? -1 - k : 2 * nz - 1 - k);
    }
    if (i < 0 || i >= nx) {
      if (g.per(0)) return V(a, wrapIdx(i, nx), j, k);
      const int im = i < 0 ? -1 - i : 2 * nx - 1 - i;
      const int f = i < 0 ? 0 : 1;
      if (g.face[f].type == NBL_BC_INLET) return 2.0 * inletComp(f, j, 1) - V(a, im, j, k);
      if (g.face[f].type == NBL_BC_OUTLET) return a[iv(std::clamp(i, 0, nx - 1), j, k)];
      return 2.0 * g.face[f].wall_velocity[1] - V(a, im, j, k);
    }
    if (j < 0 || j > ny) {
      if (g.per(1)) return V(a, i, wrapIdx(j, ny), k);
      const int jb = j < 0 ? 0 : ny;
      const int f = j < 0 ? 2 : 3;
      if (g.face[f].type == NBL_BC_OUTLET) return a[iv(i, jb, k)];
      return 2.0 * a[iv(i, jb, k)] - V(a, i, 2 * jb - j, k);
    }
    return a[iv(i, j, k)];
  }
  double W(const Vec& a, int i, int j, int k) const {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (j < 0 || j >= ny) {
      if (g.per(1)) return W(a, i, wrapIdx(j, ny), k);
      const int f = j < 0 ? 2 : 3;
      return 2.0 * g.face[f].wall_velocity[2] - W(a, i, j < 0 ? -1 - j : 2 * ny - 1 - j, k);
    }
    if (i < 0 || i >= nx) {
      if (g.per(0)) return W(a, wrapIdx(i, nx), j, k);
      const int im = i < 0 ? -1 - i : 2 * nx - 1 - i;
      const int f = i < 0 ? 0 : 1;
      if (g.face[f].type == NBL_BC_OUTLET) return a[iw(std::clamp(i, 0, nx - 1), j, k)];
      return 2.0 * g.face[f].wall_velocity[2] - W(a, im, j, k);
    }
    if (k < 0 || k > nz) {
      if (g.per(2)) return W(a, i, j, wrapIdx(k, nz));
      const int kb = k < 0 ? 0 : nz;
      return 2.0 * a[iw(i, j, kb)] - W(a, i, j, 2 * kb - k);
    }
    return a[iw(i, j, k)];
  }

  // Boundary velocities (ns_solver.cpp:221-309).
  void bc(Vec& uu, Vec& vv, Vec& ww, bool outflow) const {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int s = 0; s < 2; ++s) {
          const int i = s == 0 ? 0 : nx;
          switch (g.face[s].type) {
            case NBL_BC_WALL: uu[iu(i, j, k)] = 0.0; break;
            case NBL_BC_INLET: uu[iu(i, j, k)] = inletComp(s, j, 0); break;
            case NBL_BC_OUTLET:
              if (outflow) uu[iu(i, j, k)] = uu[iu(s == 0 ? 1 : nx - 1, j, k)];
              break;
            default: break;
          }
        }
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i)
        for (int s = 0; s < 2; ++s) {
          const int f = 2 + s, j = s == 0 ? 0 : ny;
          switch (g.face[f].type) {
            case NBL_BC_WALL: vv[iv(i, j, k)] = 0.0; break;
            case NBL_BC_INLET: vv[iv(i, j, k)] = inletComp(f, i, 1); break;
            case NBL_BC_OUTLET:
              if (outflow) vv[iv(i, j, k)] = vv[iv(i, s == 0 ? 1 : ny - 1, k)];
              break;
            default: break;
          }
        }
    if (nz > 1) {
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          for (int s = 0; s < 2; ++s) {
            const int f = 4 + s, k = s == 0 ? 0 : nz;
            switch (g.face[f].type) {
              case NBL_BC_WALL: ww[iw(i, j, k)] = 0.0; break;
              case NBL_BC_INLET: ww[iw(i, j, k)] = inletComp(f, i, 2); break;
              case NBL_BC_OUTLET:
                if (outflow) ww[iw(i, j, k)] = ww[iw(i, j, s == 0 ? 1 : nz - 1)];
                break;
              default: break;
            }
          }
    }
    if (!anySolid) return;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i)
          if (g.solidAt(i - 1, j, k) || g.solidAt(i, j, k)) uu[iu(i, j, k)] = 0.0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i)
          if (g.solidAt(i, j - 1, k) || g.solidAt(i, j, k)) vv[iv(i, j, k)] = 0.0;
    if (nz > 1)
      for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i)
            if (g.solidAt(i, j, k - 1) || g.solidAt(i, j, k)) ww[iw(i, j, k)] = 0.0;
  }

  // Momentum right-hand side (ns_solver.cpp:311-473).
  void rhs(const Vec& a, const Vec& b, const Vec& c, Vec& ru, Vec& rv, Vec& rw) const {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    const bool flat = nz == 1;
    const int sc = cfg.convection;
    ru.assign(a.size(), 0.0);
    rv.assign(b.size(), 0.0);
    rw.assign(c.size(), 0.0);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 1; i < nx; ++i) {
          if (g.solidAt(i - 1, j, k) || g.solidAt(i, j, k)) continue;
          const double aL = 0.5 * (U(a, i - 1, j, k) + U(a, i, j, k));
          const double aR = 0.5 * (U(a, i, j, k) + U(a, i + 1, j, k));
          const double qL = faceVal(sc, aL, U(a, i - 3, j, k), U(a, i - 2, j, k), U(a, i - 1, j, k),
                                    U(a, i, j, k), U(a, i + 1, j, k), U(a, i + 2, j, k));
          const double qR = faceVal(sc, aR, U(a, i - 2, j, k), U(a, i - 1, j, k), U(a, i, j, k),
                                    U(a, i + 1, j, k), U(a, i + 2, j, k), U(a, i + 3, j, k));
          const double fx = (aR * qR - aL * qL) / dx;
          const double vT = 0.5 * (V(b, i - 1, j + 1, k) + V(b, i, j + 1, k));
          const double vB = 0.5 * (V(b, i - 1, j, k) + V(b, i, j, k));
          const double qT = faceVal(sc, vT, U(a, i, j - 2, k), U(a, i, j - 1, k), U(a, i, j, k),
                                    U(a, i, j + 1, k), U(a, i, j + 2, k), U(a, i, j + 3, k));
          const double qB = faceVal(sc, vB, U(a, i, j - 3, k), U(a, i, j - 2, k), U(a, i, j - 1, k),
                                    U(a, i, j, k), U(a, i, j + 1, k), U(a, i, j + 2, k));
          const double fy = (vT * qT - vB * qB) / dy;
          double fz = 0.0;
          if (!flat) {
            const double wT = 0.5 * (W(c, i - 1, j, k + 1) + W(c, i, j, k + 1));
            const double wB = 0.5 * (W(c, i - 1, j, k) + W(c, i, j, k));
            const double zT = faceVal(sc, wT, U(a, i, j, k - 2), U(a, i, j, k - 1), U(a, i, j, k),
                                      U(a, i, j, k + 1), U(a, i, j, k + 2), U(a, i, j, k + 3));
            const double zB = faceVal(sc, wB, U(a, i, j, k - 3), U(a, i, j, k - 2), U(a, i, j, k - 1),
                                      U(a, i, j, k), U(a, i, j, k + 1), U(a, i, j, k + 2));
            fz = (wT * zT - wB * zB) / dz;
          }
          double lap = (U(a, i + 1, j, k) - 2 * U(a, i, j, k) + U(a, i - 1, j, k)) / (dx * dx) +
                       (U(a, i, j + 1, k) - 2 * U(a, i, j, k) + U(a, i, j - 1, k)) / (dy * dy);
          if (!flat) lap += (U(a, i, j, k + 1) - 2 * U(a, i, j, k) + U(a, i, j, k - 1)) / (dz * dz);
          ru[iu(i, j, k)] = -(fx + fy + fz) + nu * lap + cfg.body_force[0];
        }
    for (int k = 0; k < nz; ++k)
      for (int j = 1; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          if (g.solidAt(i, j - 1, k) || g.solidAt(i, j, k)) continue;
          const double bB = 0.5 * (V(b, i, j - 1, k) + V(b, i, j, k));
          const double bT = 0.5 * (V(b, i, j, k) + V(b, i, j + 1, k));
          const double qB = faceVal(sc, bB, V(b, i, j - 3, k), V(b, i, j - 2, k), V(b, i, j - 1, k),
                                    V(b, i, j, k), V(b, i, j + 1, k), V(b, i, j + 2, k));
          const double qT = faceVal(sc, bT, V(b, i, j - 2, k), V(b, i, j - 1, k), V(b, i, j, k),
                                    V(b, i, j + 1, k), V(b, i, j + 2, k), V(b, i, j + 3, k));
          const double fy = (bT * qT - bB * qB) / dy;
          const double uR = 0.5 * (U(a, i + 1, j - 1, k) + U(a, i + 1, j, k));
          const double uL = 0.5 * (U(a, i, j - 1, k) + U(a, i, j, k));
          const double qR = faceVal(sc, uR, V(b, i - 2, j, k), V(b, i - 1, j, k), V(b, i, j, k),
                                    V(b, i + 1, j, k), V(b, i + 2, j, k), V(b, i + 3, j, k));
          const double qL = faceVal(sc, uL, V(b, i - 3, j, k), V(b, i - 2, j, k), V(b, i - 1, j, k),
                                    V(b, i, j, k), V(b, i + 1, j, k), V(b, i + 2, j, k));
          const double fx = (uR * qR - uL * qL) / dx;
          double fz = 0.0;
          if (!flat) {
            const double wR = 0.5 * (W(c, i, j - 1, k + 1) + W(c, i, j, k + 1));
            const double wL = 0.5 * (W(c, i, j - 1, k) + W(c, i, j, k));
            const double zR = faceVal(sc, wR, V(b, i, j, k - 2), V(b, i, j, k - 1), V(b, i, j, k),
                                      V(b, i, j, k + 1), V(b, i, j, k + 2), V(b, i, j, k + 3));
            const double zL = faceVal(sc, wL, V(b, i, j, k - 3), V(b, i, j, k - 2), V(b, i, j, k - 1),
                                      V(b, i, j, k), V(b, i, j, k + 1), V(b, i, j, k + 2));
            fz = (wR * zR - wL * zL) / dz;
          }
          double lap = (V(b, i + 1, j, k) - 2 * V(b, i, j, k) + V(b, i - 1, j, k)) / (dx * dx) +
                       (V(b, i, j + 1, k) - 2 * V(b, i, j, k) + V(b, i, j - 1, k)) / (dy * dy);
          if (!flat) lap += (V(b, i, j, k + 1) - 2 * V(b, i, j, k) + V(b, i, j, k - 1)) / (dz * dz);
          rv[iv(i, j, k)] = -(fx + fy + fz) + nu * lap + cfg.body_force[1];
        }
    if (flat) return;
    for (int k = 1; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          if (g.solidAt(i, j, k - 1) || g.solidAt(i, j, k)) continue;
          const double cB = 0.5 * (W(c, i, j, k - 1) + W(c, i, j, k));
          const double cT = 0.5 * (W(c, i, j, k) + W(c, i, j, k + 1));
          const double qB = faceVal(sc, cB, W(c, i, j, k - 3), W(c, i, j, k - 2), W(c, i, j, k - 1),
                                    W(c, i, j, k), W(c, i, j, k + 1), W(c, i, j, k + 2));
          const double qT = faceVal(sc, cT, W(c, i, j, k - 2), W(c, i, j, k - 1), W(c, i, j, k),
                                    W(c, i, j, k + 1), W(c, i, j, k + 2), W(c, i, j, k + 3));
          const double fz = (cT * qT - cB * qB) / dz;
          const double uR = 0.5 * (U(a, i + 1, j, k - 1) + U(a, i + 1, j, k));
          const double uL = 0.5 * (U(a, i, j, k - 1) + U(a, i, j, k));
          const double xR = faceVal(sc, uR, W(c, i - 2, j, k), W(c, i - 1, j, k), W(c, i, j, k),
                                    W(c, i + 1, j, k), W(c, i + 2, j, k), W(c, i + 3, j, k));
          const double xL = faceVal(sc, uL, W(c, i - 3, j, k), W(c, i - 2, j, k), W(c, i - 1, j, k),
                                    W(c, i, j, k), W(c, i + 1, j, k), W(c, i + 2, j, k));
          const double fx = (uR * xR - uL * xL) / dx;
          const double vR = 0.5 * (V(b, i, j + 1, k - 1) + V(b, i, j + 1, k));
          const double vL = 0.5 * (V(b, i, j, k - 1) + V(b, i, j, k));
          const double yR = faceVal(sc, vR, W(c, i, j - 2, k), W(c, i, j - 1, k), W(c, i, j, k),
                                    W(c, i, j + 1, k), W(c, i, j + 2, k), W(c, i, j + 3, k));
          const double yL = faceVal(sc, vL, W(c, i, j - 3, k), W(c, i, j - 2, k), W(c, i, j - 1, k),
                                    W(c, i, j, k), W(c, i, j + 1, k), W(c, i, j + 2, k));
          const double fy = (vR * yR - vL * yL) / dy;
          const double lap = (W(c, i + 1, j, k) - 2 * W(c, i, j, k) + W(c, i - 1, j, k)) / (dx * dx) +
                             (W(c, i, j + 1, k) - 2 * W(c, i, j, k) + W(c, i, j - 1, k)) / (dy * dy) +
                             (W(c, i, j, k + 1) - 2 * W(c, i, j, k) + W(c, i, j, k - 1)) / (dz * dz);
          rw[iw(i, j, k)] = -(fx + fy + fz) + nu * lap + cfg.body_force[2];
        }
  }

  double divAt(const Vec& a, const Vec& b, const Vec& c, int i, int j, int k) const {
    double d = (a[iu(i + 1, j, k)] - a[iu(i, j, k)]) / g.dx + (b[iv(i, j + 1, k)] - b[iv(i, j, k)]) / g.dy;
    if (g.nz > 1) d += (c[iw(i, j, k + 1)] - c[iw(i, j, k)]) / g.dz;
    return d;
  }

  // Pressure projection (ns_solver.cpp:475-579).
  void project(Vec& a, Vec& b, Vec& c, double dt, int& iters, double& res) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    bc(a, b, c, true);
    Vec rhsv(g.cells(), 0.0);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          if (g.solidAt(i, j, k)) continue;
          rhsv[g.c(i, j, k)] = -divAt(a, b, c, i, j, k);
        }
    Vec phi;
    const nbl_poisson_report rep = mg.solve(g, rhsv, phi);
    iters = rep.iterations;
    res = rep.residual;
    // Ghost value of phi: periodic wrap, Neumann copy, outlet Dirichlet (-phi).
    auto ph = [&](int i, int j, int k) {
      int q[3] = {i, j, k};
      double s = 1.0;
      for (int ax = 0; ax < 3; ++ax) {
        const int n = g.n(ax);
        if (q[ax] >= 0 && q[ax] < n) continue;
        if (g.per(ax)) {
          q[ax] = wrapIdx(q[ax], n);
          continue;
        }
        const int f = 2 * ax + (q[ax] < 0 ? 0 : 1);
        q[ax] = std::clamp(q[ax], 0, n - 1);
        s *= g.face[f].type == NBL_BC_OUTLET ? -1.0 : 1.0;
      }
      return s * phi[g.c(q[0], q[1], q[2])];
    };
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
          if (g.solidAt(i - 1, j, k) || g.solidAt(i, j, k)) continue;
          a[iu(i, j, k)] -= (ph(i, j, k) - ph(i - 1, j, k)) / g.dx;
        }
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
          if (g.solidAt(i, j - 1, k) || g.solidAt(i, j, k)) continue;
          b[iv(i, j, k)] -= (ph(i, j, k) - ph(i, j - 1, k)) / g.dy;
        }
    if (nz > 1)
      for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            if (g.solidAt(i, j, k - 1) || g.solidAt(i, j, k)) continue;
            c[iw(i, j, k)] -= (ph(i, j, k) - ph(i, j, k - 1)) / g.dz;
          }
    bc(a, b, c, false);
    for (size_t q = 0; q < p.size(); ++q) p[q] = phi[q] / dt;
  }

  // CFL time step (ns_solver.cpp:581-603).
  double cflDt() const {
    const bool flat = g.nz == 1;
    const double visc = 2.0 * nu * (1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy) + (flat ? 0.0 : 1.0 / (g.dz * g.dz)));
    double m = visc + 1e-300;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          if (g.solidAt(i, j, k)) continue;
          const double a = std::abs(0.5 * (u[iu(i, j, k)] + u[iu(i + 1, j, k)]));
          const double b = std::abs(0.5 * (v[iv(i, j, k)] + v[iv(i, j + 1, k)]));
          const double c = flat ? 0.0 : std::abs(0.5 * (w[iw(i, j, k)] + w[iw(i, j, k + 1)]));
          const double s = a / g.dx + b / g.dy + (flat ? 0.0 : c / g.dz) + visc;
          m = std::max(m, s);
        }
    return cfg.cfl / m;
  }

  double divL2() const {
    double s = 0.0;
    size_t n = 0;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          if (g.solidAt(i, j, k)) continue;
          const double d = divAt(u, v, w, i, j, k);
          s += d * d;
          ++n;
        }
    return n ? std::sqrt(s / double(n)) : 0.0;
  }
  double divMax() const {
    double m = 0.0;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          if (g.solidAt(i, j, k)) continue;
          m = std::max(m, std::abs(divAt(u, v, w, i, j, k)));
        }
    return m;
  }
  double massErr() const {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double ax = g.dy * g.dz, ay = g.dx * g.dz, az = g.dx * g.dy;
    double net = 0.0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) net += (u[iu(0, j, k)] - u[iu(nx, j, k)]) * ax;
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i) net += (v[iv(i, 0, k)] - v[iv(i, ny, k)]) * ay;
    if (nz > 1)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) net += (w[iw(i, j, 0)] - w[iw(i, j, nz)]) * az;
    return std::abs(net) / (cfg.velocity_scale * (ny * ay) + 1e-300);
  }
  double cU(int i, int j, int k) const { return 0.5 * (u[iu(i, j, k)] + u[iu(i + 1, j, k)]); }
  double cV(int i, int j, int k) const { return 0.5 * (v[iv(i, j, k)] + v[iv(i, j + 1, k)]); }
  double cW(int i, int j, int k) const { return g.nz == 1 ? 0.0 : 0.5 * (w[iw(i, j, k)] + w[iw(i, j, k + 1)]); }

  // Drag/lift from fluid cells next to solid cells (ns_solver.cpp:703-763).
  void forces(double& cd, double& cl) const {
    cd = cl = 0.0;
    if (!anySolid) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double aX = g.dy * g.dz, aY = g.dx * g.dz, aZ = g.dx * g.dy;
    const double mu = nu * cfg.density;
    double fx = 0.0, fy = 0.0;
    auto add = [&](int i, int j, int k, int si, int sj, int sk, double nx_, double ny_, double area,
                   int axis, double dn) {
      if (!g.solidAt(si, sj, sk) || g.solidAt(i, j, k)) return;
      if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return;
      const double pf = p[g.c(i, j, k)];
      fx += -pf * nx_ * area;
      fy += -pf * ny_ * area;
      const double a = cU(i, j, k), b = cV(i, j, k);
      if (axis != 0) fx += mu * a / (0.5 * dn) * area;
      if (axis != 1) fy += mu * b / (0.5 * dn) * area;
    };
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          add(i, j, k, i - 1, j, k, 1.0, 0.0, aX, 0, g.dx);
          add(i, j, k, i + 1, j, k, -1.0, 0.0, aX, 0, g.dx);
          add(i, j, k, i, j - 1, k, 0.0, 1.0, aY, 1, g.dy);
          add(i, j, k, i, j + 1, k, 0.0, -1.0, aY, 1, g.dy);
          if (nz > 1) {
            add(i, j, k, i, j, k - 1, 0.0, 0.0, aZ, 2, g.dz);
            add(i, j, k, i, j, k + 1, 0.0, 0.0, aZ, 2, g.dz);
          }
        }
    double aref = cfg.ref_area;
    if (aref <= 0.0) aref = cfg.ref_length * cfg.ref_length;
    const double q = 0.5 * cfg.density * cfg.velocity_scale * cfg.velocity_scale * aref + 1e-300;
    cd = fx / q;
    cl = fy / q;
  }

  // SSP-RK3 step with a projection per stage (ns_solver.cpp:767-824).
  nbl_step_report advance(double dt) {
    const Vec u0 = u, v0 = v, w0 = w;
    Vec ru, rv, rw;
    int i1 = 0, i2 = 0, i3 = 0;
    double r1 = 0, r2 = 0, r3 = 0;
    bc(u, v, w, true);
    rhs(u, v, w, ru, rv, rw);
    for (size_t m = 0; m < u.size(); ++m) u[m] = u0[m] + dt * ru[m];
    for (size_t m = 0; m < v.size(); ++m) v[m] = v0[m] + dt * rv[m];
    for (size_t m = 0; m < w.size(); ++m) w[m] = w0[m] + dt * rw[m];
    project(u, v, w, dt, i1, r1);
    rhs(u, v, w, ru, rv, rw);
    for (size_t m = 0; m < u.size(); ++m) u[m] = 0.75 * u0[m] + 0.25 * (u[m] + dt * ru[m]);
    for (size_t m = 0; m < v.size(); ++m) v[m] = 0.75 * v0[m] + 0.25 * (v[m] + dt * rv[m]);
    for (size_t m = 0; m < w.size(); ++m) w[m] = 0.75 * w0[m] + 0.25 * (w[m] + dt * rw[m]);
    project(u, v, w, dt, i2, r2);
    rhs(u, v, w, ru, rv, rw);
    const double c1 = 1.0 / 3.0, c2 = 2.0 / 3.0;
    for (size_t m = 0; m < u.size(); ++m) u[m] = c1 * u0[m] + c2 * (u[m] + dt * ru[m]);
    for (size_t m = 0; m < v.size(); ++m) v[m] = c1 * v0[m] + c2 * (v[m] + dt * rv[m]);
    for (size_t m = 0; m < w.size(); ++m) w[m] = c1 * w0[m] + c2 * (w[m] + dt * rw[m]);
    project(u, v, w, dt, i3, r3);
    double du = 0.0;
    size_t n = 0;
    for (size_t m = 0; m < u.size(); ++m) { du += (u[m] - u0[m]) * (u[m] - u0[m]); ++n; }
    for (size_t m = 0; m < v.size(); ++m) { du += (v[m] - v0[m]) * (v[m] - v0[m]); ++n; }
    for (size_t m = 0; m < w.size(); ++m) { du += (w[m] - w0[m]) * (w[m] - w0[m]); ++n; }
    const double mom = std::sqrt(du / double(n ? n : 1)) / (dt * cfg.velocity_scale + 1e-300);
    nbl_step_report r{};
    stepN += 1;
    t += dt;
    dtLast = dt;
    r.step = stepN;
    r.t = t;
    r.dt = dt;
    r.cfl = cfg.cfl;
    r.momentum_residual = mom;
    r.continuity_residual = divL2();
    r.divergence_max = divMax();
    r.mass_error = massErr();
    r.poisson_iters = i1 + i2 + i3;
    r.poisson_residual = std::max({r1, r2, r3});
    forces(r.cd, r.cl);
    r.cell_count = g.cells();
    r.accepted = (r.continuity_residual < cfg.tol_continuity && r.mass_error < cfg.tol_mass) ? 1 : 0;
    return r;
  }
};

// ---------------------------------------------------------------------------
// Morton keys (core/include/nabla/mesh/morton.hpp:15-48): 21 bits per axis,
// x at bit 3i, y at 3i+1, z at 3i+2.  Written as plain bit loops.
// ---------------------------------------------------------------------------
static uint64_t spread(uint32_t a) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) r |= uint64_t((a >> b) & 1u) << (3 * b);
  return r;
}
static uint32_t gather(uint64_t m) {
  uint32_t r = 0;
  for (int b = 0; b < 21; ++b) r |= uint32_t((m >> (3 * b)) & 1u) << b;
  return r;
}
static uint64_t menc(uint32_t x, uint32_t y, uint32_t z) { return spread(x) | (spread(y) << 1) | (spread(z) << 2); }

// ---------------------------------------------------------------------------
// Linear octree (core/src/mesh/octree.cpp).  Leaves kept in the reference's
// insertion / swap-remove order so that serial sums run in the same order.
// ---------------------------------------------------------------------------
struct Oct {
  double org[3], ext[3];
  int maxL;
  double fine[3];
  std::vector<uint64_t> key;
  std::vector<uint8_t> lev, msk;
  std::unordered_map<uint64_t, uint32_t> at;

  Oct(const double o[3], const double e[3], int ml) : maxL(ml) {
    if (ml < 1 || ml > 21) throw std::invalid_argument("maxLevel");
    const double inv = 1.0 / double(1u << ml);
    for (int a = 0; a < 3; ++a) {
      org[a] = o[a];
      ext[a] = e[a];
      fine[a] = e[a] * inv;
    }
    push(0, 0, 0);
  }
  uint32_t side(int l) const { return 1u << (maxL - l); }
  size_t push(uint64_t k, int l, uint8_t m) {
    const size_t i = key.size();
    key.push_back(k);
    lev.push_back(uint8_t(l));
    msk.push_back(m);
    at[k] = uint32_t(i);
    return i;
  }
  void drop(size_t i) {  // swap-remove (octree.cpp:171-196)
    const uint64_t gone = key[i];
    const size_t last = key.size() - 1;
    if (i != last) {
      key[i] = key[last];
      lev[i] = lev[last];
      msk[i] = msk[last];
      at[key[i]] = uint32_t(i);
    }
    key.pop_back();
    lev.pop_back();
    msk.pop_back();
    at.erase(gone);
  }
  double center(size_t i, int axis) const {
    const uint32_t a = gather(key[i] >> axis);
    const double units = double(a) + 0.5 * double(side(lev[i]));
    return org[axis] + units * fine[axis];
  }
  double size(size_t i, int axis) const { return double(side(lev[i])) * fine[axis]; }
  long find(uint64_t k) const {
    auto it = at.find(k);
    return it == at.end() ? -1 : long(it->second);
  }
  // Same-or-coarser leaf across face `dir` (octree.cpp:201-227); -1 if none.
  long coarseNb(const uint32_t a[3], int l, int dir, int& nl) const {
    const int axis = dir / 2;
    int64_t q[3] = {a[0], a[1], a[2]};
    q[axis] += (dir & 1 ? 1 : -1) * int64_t(side(l));
    if (q[axis] < 0 || q[axis] >= (int64_t(1) << maxL)) return -1;
    for (int L = l; L >= 0; --L) {
      const uint32_t m = ~(side(L) - 1u);
      const long hit = find(menc(uint32_t(q[0]) & m, uint32_t(q[1]) & m, uint32_t(q[2]) & m));
      if (hit >= 0 && lev[size_t(hit)] == L) {
        nl = L;
        return hit;
      }
    }
    return -1;
  }
  // All leaves across face `dir` (octree.cpp:229-277).
  std::vector<size_t> nbrs(size_t i, int dir) const {
    uint32_t a[3] = {gather(key[i]), gather(key[i] >> 1), gather(key[i] >> 2)};
    const int l = lev[i];
    const int axis = dir / 2;
    const bool plus = dir & 1;
    int nl = 0;
    int64_t q[3] = {a[0], a[1], a[2]};
    q[axis] += (plus ? 1 : -1) * int64_t(side(l));
    if (q[axis] < 0 || q[axis] >= (int64_t(1) << maxL)) return {};
    const long c = coarseNb(a, l, dir, nl);
    if (c >= 0) return {size_t(c)};
    std::vector<size_t> out;
    if (l + 1 > maxL) return out;
    const uint32_t cs = side(l + 1);
    const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
    const uint32_t n[3] = {uint32_t(q[0]), uint32_t(q[1]), uint32_t(q[2])};
    const uint32_t base = plus ? n[axis] : n[axis] + side(l) - cs;
    for (int k1 = 0; k1 < 2; ++k1)
      for (int k2 = 0; k2 < 2; ++k2) {
        uint32_t b[3];
        b[axis] = base;
        b[t1] = n[t1] + uint32_t(k1) * cs;
        b[t2] = n[t2] + uint32_t(k2) * cs;
        const long hit = find(menc(b[0], b[1], b[2]));
        if (hit >= 0 && lev[size_t(hit)] == l + 1) out.push_back(size_t(hit));
      }
    return out;
  }
  void split(size_t i) {  // octree.cpp:288-323 (fields omitted: not on the path)
    const uint64_t k = key[i];
    const int l = lev[i];
    const uint8_t m = msk[i];
    const uint32_t a[3] = {gather(k), gather(k >> 1), gather(k >> 2)};
    const uint32_t cs = side(l + 1);
    drop(i);
    for (int c = 0; c < 8; ++c)
      push(menc(a[0] + uint32_t(c & 1) * cs, a[1] + uint32_t((c >> 1) & 1) * cs, a[2] + uint32_t((c >> 2) & 1) * cs),
           l + 1, m);
  }
  void refineKey(uint64_t k) {  // balanced refinement (octree.cpp:325-356)
    if (find(k) < 0) throw std::runtime_error("refine: cell not found");
    for (;;) {
      const size_t i = size_t(find(k));
      const int l = lev[i];
      if (l >= maxL) throw std::runtime_error("refine: cell already at maxLevel");
      const uint32_t a[3] = {gather(k), gather(k >> 1), gather(k >> 2)};
      bool again = false;
      for (int d = 0; d < 6; ++d) {
        int nl = 0;
        const long nb = coarseNb(a, l, d, nl);
        if (nb >= 0 && nl < l) {
          refineKey(key[size_t(nb)]);
          again = true;
          break;
        }
      }
      if (!again) break;
    }
    split(size_t(find(k)));
  }
  void refineUniform(int times) {
    for (int t = 0; t < times; ++t) {
      const std::vector<uint64_t> ks = key;
      for (uint64_t k : ks) {
        const long i = find(k);
        if (i >= 0) split(size_t(i));
      }
    }
  }
  void refineBox(const double lo[3], const double hi[3], int times) {
    for (int t = 0; t < times; ++t) {
      std::vector<uint64_t> ks;
      for (size_t i = 0; i < key.size(); ++i) {
        const double cx = center(i, 0), cy = center(i, 1), cz = center(i, 2);
        if (cx >= lo[0] && cx <= hi[0] && cy >= lo[1] && cy <= hi[1] && cz >= lo[2] && cz <= hi[2])
          ks.push_back(key[i]);
      }
      for (uint64_t k : ks)
        if (find(k) >= 0) refineKey(k);
    }
  }
  bool balanced() const {
    for (size_t i = 0; i < key.size(); ++i) {
      const uint32_t a[3] = {gather(key[i]), gather(key[i] >> 1), gather(key[i] >> 2)};
      for (int d = 0; d < 6; ++d) {
        int nl = 0;
        if (coarseNb(a, lev[i], d, nl) >= 0 && lev[i] - nl > 1) return false;
      }
    }
    return true;
  }
  bool fluid(size_t i) const { return msk[i] == 0; }
  // area of the finer cell / centre distance (octree_poisson.cpp:24-32)
  double coeff(size_t i, size_t j, int axis) const {
    const int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
    const size_t f = lev[j] > lev[i] ? j : i;
    const double area = size(f, t1) * size(f, t2);
    const double dist = std::abs(center(i, axis) - center(j, axis));
    return dist > 0.0 ? area / dist : 0.0;
  }
  void apply(const Vec& x, Vec& out) const {  // octree_poisson.cpp:64-85
    const size_t n = key.size();
    out.assign(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
      if (!fluid(i)) {
        out[i] = x[i];
        continue;
      }
      double acc = 0.0;
      for (int d = 0; d < 6; ++d)
        for (size_t j : nbrs(i, d)) {
          if (!fluid(j)) continue;
          acc += coeff(i, j, d / 2) * (x[i] - x[j]);
        }
      out[i] = acc;
    }
  }
  void diag(Vec& dg) const {  // octree_poisson.cpp:87-105
    const size_t n = key.size();
    dg.assign(n, 1.0);
    for (size_t i = 0; i < n; ++i) {
      if (!fluid(i)) continue;
      double d = 0.0;
      for (int dd = 0; dd < 6; ++dd)
        for (size_t j : nbrs(i, dd))
          if (fluid(j)) d += coeff(i, j, dd / 2);
      dg[i] = d > 0.0 ? d : 1.0;
    }
  }
  // First-order conservative upwind step of a cell scalar with cell-centred
  // velocities (transport.cpp:24-60): closed boundaries, shared face area =
  // the smaller transverse area, outward velocity = the face average; returns
  // the integral sum_i phi_i vol_i (transport.cpp:9-17).
  double transport(const Vec& u, const Vec& v, const Vec& w, Vec& phi, double dt) const {
    const size_t n = key.size();
    Vec dphi(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
      const double szi[3] = {size(i, 0), size(i, 1), size(i, 2)};
      const double vol = szi[0] * szi[1] * szi[2];
      double netOut = 0.0;
      for (int dir = 0; dir < 6; ++dir) {
        const int axis = dir / 2;
        const double sign = (dir & 1) ? 1.0 : -1.0;
        for (size_t nb : nbrs(i, dir)) {
          const double szn[3] = {size(nb, 0), size(nb, 1), size(nb, 2)};
          const double areaI = axis == 0 ? szi[1] * szi[2] : axis == 1 ? szi[0] * szi[2] : szi[0] * szi[1];
          const double areaN = axis == 0 ? szn[1] * szn[2] : axis == 1 ? szn[0] * szn[2] : szn[0] * szn[1];
          const double area = std::min(areaI, areaN);
          const Vec& vel = axis == 0 ? u : axis == 1 ? v : w;
          const double vn = sign * 0.5 * (vel[i] + vel[nb]);
          const double up = vn >= 0.0 ? phi[i] : phi[nb];
          netOut += vn * area * up;
        }
      }
      dphi[i] = -dt / vol * netOut;
    }
    for (size_t i = 0; i < n; ++i) phi[i] += dphi[i];
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += phi[i] * (size(i, 0) * size(i, 1) * size(i, 2));
    return sum;
  }
  // Gradient of f at leaf i across level jumps (octree.cpp:491-539): per axis
  // the mean value and mean centre of the leaves across the + and - faces (the
  // leaf itself stands in for a missing side), differenced.
  std::array<double, 3> gradient(const Vec& f, size_t i) const {
    std::array<double, 3> g{0.0, 0.0, 0.0};
    for (int a = 0; a < 3; ++a) {
      double val[2] = {f[i], f[i]}, pos[2] = {center(i, a), center(i, a)};
      bool has[2] = {false, false};
      for (int side = 0; side < 2; ++side) {  // 0: + face (2a+1), 1: - face (2a)
        const auto nb = nbrs(i, side == 0 ? 2 * a + 1 : 2 * a);
        if (nb.empty()) continue;
        double sv = 0.0, sp = 0.0;
        for (size_t j : nb) {
          sv += f[j];
          sp += center(j, a);
        }
        const double inv = 1.0 / double(nb.size());
        val[side] = sv * inv;
        pos[side] = sp * inv;
        has[side] = true;
      }
      if (!has[0] && !has[1]) continue;
      const double d = pos[0] - pos[1];
      g[a] = d != 0.0 ? (val[0] - val[1]) / d : 0.0;
    }
    return g;
  }
  // PhysicalSensorEvaluator::evaluate (adaptive_core.cpp:73-176): out[f*n+i]
  // for f = vorticity magnitude, strain magnitude, Q, |grad p|, |grad u|, y+,
  // wall shear, truncation estimate; flags[f*n+i] for near-wall, reverse flow,
  // residual stagnation; speedPrev / residPrev are the carried state.
  void sensors(const Vec& U, const Vec& V, const Vec& W, const Vec& P, double nu, Vec& speedPrev, Vec& residPrev,
               double* out, uint8_t* flags) const {
    const size_t n = key.size();
    for (size_t i = 0; i < n; ++i) {
      const auto du = gradient(U, i), dv = gradient(V, i), dw = gradient(W, i), dp = gradient(P, i);
      const double ox = dw[1] - dv[2], oy = du[2] - dw[0], oz = dv[0] - du[1];
      const double vort = std::sqrt(ox * ox + oy * oy + oz * oz);
      const double sxx = du[0], syy = dv[1], szz = dw[2];
      const double sxy = 0.5 * (du[1] + dv[0]), sxz = 0.5 * (du[2] + dw[0]), syz = 0.5 * (dv[2] + dw[1]);
      const double ss = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz);
      const double rxy = 0.5 * (du[1] - dv[0]), rxz = 0.5 * (du[2] - dw[0]), ryz = 0.5 * (dv[2] - dw[1]);
      const double oo = 2.0 * (rxy * rxy + rxz * rxz + ryz * ryz);
      const double gu = std::sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2] + dv[0] * dv[0] + dv[1] * dv[1] +
                                  dv[2] * dv[2] + dw[0] * dw[0] + dw[1] * dw[1] + dw[2] * dw[2]);
      const double gp = std::sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
      // near a wall: no leaf across a y face (floor, ceiling) or a solid neighbour
      bool wall = nbrs(i, 2).empty() || nbrs(i, 3).empty();
      for (int d = 0; d < 6 && !wall; ++d)
        for (size_t j : nbrs(i, d))
          if (msk[j] == 2) wall = true;
      const double speed = std::sqrt(U[i] * U[i] + V[i] * V[i] + W[i] * W[i]);
      double yplus = -1.0, tau = 0.0;
      if (wall) {
        const double yw = 0.5 * size(i, 1);
        const double ut = std::sqrt(U[i] * U[i] + W[i] * W[i]);
        tau = nu * ut / (yw + 1e-300);
        yplus = std::sqrt(tau) * yw / (nu + 1e-300);
      }
      double msum = 0.0;
      int cnt = 0;
      for (int d = 0; d < 6; ++d)
        for (size_t j : nbrs(i, d)) {
          msum += std::sqrt(U[j] * U[j] + V[j] * V[j] + W[j] * W[j]);
          ++cnt;
        }
      const double vals[8] = {vort, std::sqrt(2.0 * ss), 0.5 * (oo - ss), gp, gu, yplus, tau,
                              cnt ? std::abs(speed - msum / cnt) : 0.0};
      for (int f = 0; f < 8; ++f) out[size_t(f) * n + i] = vals[f];
      flags[i] = wall ? 1 : 0;
      flags[n + i] = (wall && U[i] < 0.0) ? 1 : 0;
    }
    for (size_t i = 0; i < n; ++i) {  // residual stagnation: the second pass
      const double speed = std::sqrt(U[i] * U[i] + V[i] * V[i] + W[i] * W[i]);
      const double res = std::abs(speed - speedPrev[i]);
      flags[2 * n + i] = (residPrev[i] > 1e-9 && res >= 0.95 * residPrev[i]) ? 1 : 0;
      residPrev[i] = res;
      speedPrev[i] = speed;
    }
  }
  // Jacobi-PCG, always singular (octree_poisson.cpp:107-197).
  nbl_poisson_report solve(const Vec& b, Vec& x, double tol, int maxIt) const {
    const size_t n = key.size();
    x.assign(n, 0.0);
    std::vector<uint8_t> fl(n, 0);
    size_t nf = 0;
    for (size_t i = 0; i < n; ++i)
      if (fluid(i)) {
        fl[i] = 1;
        ++nf;
      }
    auto demean = [&](Vec& a) {
      if (nf == 0) return;
      double m = 0.0;
      for (size_t c = 0; c < a.size(); ++c) {
        if (fl[c]) m += a[c];
        else a[c] = 0.0;
      }
      m /= double(nf);
      for (size_t c = 0; c < a.size(); ++c)
        if (fl[c]) a[c] -= m;
    };
    Vec dg;
    diag(dg);
    Vec rhs = b;
    for (size_t c = 0; c < n; ++c)
      if (!fl[c]) rhs[c] = 0.0;
    demean(rhs);
    const double bn = std::sqrt(dotv(rhs, rhs)) + 1e-300;
    Vec r = rhs, z(n, 0.0), p(n, 0.0), Ap(n, 0.0);
    for (size_t c = 0; c < n; ++c) z[c] = fl[c] ? r[c] / dg[c] : 0.0;
    demean(z);
    p = z;
    double rz = dotv(r, z);
    nbl_poisson_report rep{0, 0.0, 0};
    int it = 0;
    for (; it < maxIt; ++it) {
      apply(p, Ap);
      const double pAp = dotv(p, Ap);
      if (std::abs(pAp) < 1e-300) break;
      const double al = rz / pAp;
      for (size_t c = 0; c < n; ++c) {
        x[c] += al * p[c];
        r[c] -= al * Ap[c];
      }
      demean(r);
      if (std::sqrt(dotv(r, r)) <= tol * bn) {
        ++it;
        rep.converged = 1;
        break;
      }
      for (size_t c = 0; c < n; ++c) z[c] = fl[c] ? r[c] / dg[c] : 0.0;
      demean(z);
      const double rz2 = dotv(r, z);
      const double be = rz2 / rz;
      for (size_t c = 0; c < n; ++c) p[c] = z[c] + be * p[c];
      rz = rz2;
    }
    demean(x);
    apply(x, Ap);
    double rr = 0.0;
    for (size_t c = 0; c < n; ++c) {
      const double e = rhs[c] - Ap[c];
      rr += e * e;
    }
    rep.iterations = it;
    rep.residual = std::sqrt(rr) / bn;
    rep.converged = (rep.converged || rep.residual <= 1e-6) ? 1 : 0;
    return rep;
  }
};


// ---- octree multigrid (no reference counterpart) -----------------------------
// Restates the device octree multigrid.  Cells of every level are boxes in
// finest-index units (anchor, log2 size per axis); level 0 = the leaves in
// Morton order.  A coarsening step raises the target log size S per axis:
// from S = the finest leaf size, the axes whose physical target size is at
// most 1.5x the smallest grow by one (semi-coarsening along the strongly
// coupled axes of anisotropic cells).  Cells no larger than S on every axis
// group by their S-aligned box; a group that covers its box exactly becomes
// one coarse cell, every other cell stays as it is.  Coarse couplings sum the
// fine face areas between two coarse cells (fluid-fluid faces only) and use
// the two-point coefficient area / centre distance (on octree cells this is
// the reference's face coefficient, octree_poisson.cpp:24-32); a coarse cell
// is fluid iff a child is fluid.  Residuals restrict by summation, corrections
// prolong by injection; V(2,2) damped Jacobi (omega 0.8, multigrid.cpp:
// 102-114); coarsest level (<= 2048 cells) solved exactly through
// (A + alpha 11^T) with the result projected onto mean-free vectors; outer
// loop = MgPoisson::solve's flexible CG in its singular form
// (multigrid.cpp:267-355).
struct OctMgOrc {
  struct Cpl {
    size_t j;
    double coef, area;
    int axis;
  };
  struct Lvl {
    std::vector<uint32_t> a[3];
    std::vector<int> l[3];
    std::vector<uint8_t> fl;
    std::vector<std::vector<Cpl>> rows;
    Vec diag;
    std::vector<size_t> up;  // parent in level l+1
    size_t n = 0, nf = 0;
    int S[3] = {0, 0, 0};
    bool fluid(size_t i) const { return fl[i] != 0; }
  };
  std::vector<Lvl> L;
  std::vector<size_t> order;  // Morton position -> caller index
  double org[3], fine[3];
  int maxL = 0;
  bool direct = false;
  std::vector<double> lu;     // LU of the regularised coarsest matrix (row-major, no pivoting needed: SPD)
  static constexpr double kOm = 0.8;

  double centre(const Lvl& v, size_t i, int ax) const {
    return org[ax] + (double(v.a[ax][i]) + 0.5 * double(uint64_t(1) << v.l[ax][i])) * fine[ax];
  }
  static void finishRows(Lvl& v) {
    v.diag.assign(v.n, 1.0);
    v.nf = 0;
    for (size_t i = 0; i < v.n; ++i) {
      if (!v.fluid(i)) continue;
      ++v.nf;
      double d = 0.0;
      for (const Cpl& c : v.rows[i]) d += c.coef;
      v.diag[i] = d > 0.0 ? d : 1.0;
    }
  }
  static double rowA(const Lvl& v, size_t i, const Vec& x) {
    double acc = 0.0;
    for (const Cpl& c : v.rows[i]) acc += c.coef * (x[i] - x[c.j]);
    return acc;
  }

  explicit OctMgOrc(const Oct& o) : maxL(o.maxL) {
    for (int q = 0; q < 3; ++q) {
      org[q] = o.org[q];
      fine[q] = o.fine[q];
    }
    const size_t n0 = o.key.size();
    order.resize(n0);
    for (size_t i = 0; i < n0; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return o.key[a] < o.key[b]; });
    std::vector<size_t> pos(n0);
    for (size_t q = 0; q < n0; ++q) pos[order[q]] = q;
    // level 0: the leaves in Morton order, couplings in the reference's order
    L.emplace_back();
    Lvl& z = L[0];
    z.n = n0;
    int lmin = 1 << 30;
    for (int q = 0; q < 3; ++q) {
      z.a[q].resize(n0);
      z.l[q].resize(n0);
    }
    z.fl.resize(n0);
    z.rows.assign(n0, {});
    for (size_t q = 0; q < n0; ++q) {
      const size_t i = order[q];
      const uint64_t k = o.key[i];
      const uint32_t an[3] = {gather(k), gather(k >> 1), gather(k >> 2)};
      for (int ax = 0; ax < 3; ++ax) {
        z.a[ax][q] = an[ax];
        z.l[ax][q] = o.maxL - o.lev[i];
      }
      lmin = std::min(lmin, o.maxL - int(o.lev[i]));
      z.fl[q] = o.fluid(i) ? 1 : 0;
      if (!o.fluid(i)) continue;
      for (int dd = 0; dd < 6; ++dd)
        for (size_t j : o.nbrs(i, dd)) {
          if (!o.fluid(j)) continue;
          const int ax = dd / 2, t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
          const size_t f = o.lev[j] > o.lev[i] ? j : i;
          z.rows[q].push_back({pos[j], o.coeff(i, j, ax), o.size(f, t1) * o.size(f, t2), ax});
        }
    }
    finishRows(z);
    for (int ax = 0; ax < 3; ++ax) z.S[ax] = lmin;
    while (L.back().n > 2048 && L.size() < 16) {
      if (!coarsen()) break;
    }
    direct = L.back().n <= 2048 && factor();
  }

  // One coarsening step from the last level; false when nothing merges up to
  // the root size.
  bool coarsen() {
    const Lvl& f = L.back();
    int S[3] = {f.S[0], f.S[1], f.S[2]};
    for (;;) {
      if (S[0] >= maxL && S[1] >= maxL && S[2] >= maxL) return false;
      double h[3], hmin = 1e300;
      for (int ax = 0; ax < 3; ++ax) {
        h[ax] = fine[ax] * double(uint64_t(1) << S[ax]);
        hmin = std::min(hmin, h[ax]);
      }
      for (int ax = 0; ax < 3; ++ax)
        if (h[ax] <= 1.5 * hmin && S[ax] < maxL) ++S[ax];
      // group the cells no larger than S by their S-aligned box
      std::map<std::array<uint32_t, 3>, std::pair<uint64_t, std::vector<size_t>>> grp;
      for (size_t i = 0; i < f.n; ++i) {
        if (f.l[0][i] > S[0] || f.l[1][i] > S[1] || f.l[2][i] > S[2]) continue;
        std::array<uint32_t, 3> t;
        for (int ax = 0; ax < 3; ++ax) t[ax] = f.a[ax][i] & ~((uint32_t(1) << S[ax]) - 1u);
        auto& g = grp[t];
        g.first += uint64_t(1) << (f.l[0][i] + f.l[1][i] + f.l[2][i]);
        g.second.push_back(i);
      }
      const uint64_t full = uint64_t(1) << (S[0] + S[1] + S[2]);
      // coarse cells keyed by anchor (Morton order)
      std::vector<std::array<uint32_t, 3>> ca;
      std::vector<std::array<int, 3>> cl;
      std::vector<size_t> tag(f.n, SIZE_MAX);
      std::map<uint64_t, size_t> byKey;
      auto add = [&](const std::array<uint32_t, 3>& an, const std::array<int, 3>& lg) {
        const uint64_t k = menc(an[0], an[1], an[2]);
        auto it = byKey.find(k);
        if (it != byKey.end()) return it->second;
        ca.push_back(an);
        cl.push_back(lg);
        byKey[k] = ca.size() - 1;
        return ca.size() - 1;
      };
      for (auto& [t, g] : grp)
        if (g.first == full && g.second.size() > 1) {
          const size_t c = add(t, {S[0], S[1], S[2]});
          for (size_t i : g.second) tag[i] = c;
        }
      for (size_t i = 0; i < f.n; ++i)
        if (tag[i] == SIZE_MAX) tag[i] = add({f.a[0][i], f.a[1][i], f.a[2][i]}, {f.l[0][i], f.l[1][i], f.l[2][i]});
      if (ca.size() == f.n) {
        if (S[0] >= maxL && S[1] >= maxL && S[2] >= maxL) return false;
        continue;  // nothing merged at this size: grow the target further
      }
      // renumber coarse cells in Morton order of their anchors
      std::vector<size_t> rank(ca.size());
      {
        size_t r = 0;
        for (auto& [k, c] : byKey) rank[c] = r++;
      }
      Lvl c;
      c.n = ca.size();
      for (int ax = 0; ax < 3; ++ax) {
        c.a[ax].resize(c.n);
        c.l[ax].resize(c.n);
        c.S[ax] = S[ax];
      }
      c.fl.assign(c.n, 0);
      for (size_t q = 0; q < ca.size(); ++q)
        for (int ax = 0; ax < 3; ++ax) {
          c.a[ax][rank[q]] = ca[q][ax];
          c.l[ax][rank[q]] = cl[q][ax];
        }
      std::vector<size_t> up(f.n);
      for (size_t i = 0; i < f.n; ++i) {
        up[i] = rank[tag[i]];
        if (f.fluid(i)) c.fl[up[i]] = 1;
      }
      // couplings: fine face areas summed per coarse pair in fine coupling order
      std::map<std::pair<size_t, size_t>, std::pair<double, int>> acc;
      for (size_t i = 0; i < f.n; ++i)
        for (const Cpl& e : f.rows[i]) {
          const size_t I = up[i], J = up[e.j];
          if (I == J) continue;
          auto it = acc.find({I, J});
          if (it == acc.end()) acc[{I, J}] = {e.area, e.axis};
          else it->second.first += e.area;
        }
      c.rows.assign(c.n, {});
      for (auto& [ij, v] : acc) {
        const auto [I, J] = ij;
        const int ax = v.second;
        const double dist = std::abs(centre(c, I, ax) - centre(c, J, ax));
        c.rows[I].push_back({J, dist > 0.0 ? v.first / dist : 0.0, v.first, ax});
      }
      finishRows(c);
      L.back().up = up;
      L.push_back(std::move(c));
      return true;
    }
  }

  // LU (Doolittle, no pivoting) of the coarsest matrix: fluid rows A + alpha
  // on the fluid block, other rows identity.
  bool factor() {
    const Lvl& v = L.back();
    const size_t n = v.n;
    lu.assign(n * n, 0.0);
    double dfl = 0.0;
    for (size_t i = 0; i < n; ++i) {
      lu[i * n + i] = v.diag[i];
      if (!v.fluid(i)) continue;
      for (const Cpl& c : v.rows[i]) lu[i * n + c.j] -= c.coef;
      dfl += v.diag[i];
    }
    if (v.nf > 0) {
      const double alpha = dfl / double(v.nf) / double(v.nf);
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          if (v.fluid(i) && v.fluid(j)) lu[i * n + j] += alpha;
    }
    for (size_t kk = 0; kk < n; ++kk) {
      const double piv = lu[kk * n + kk];
      if (!(std::abs(piv) > 1e-12)) return false;
      for (size_t i = kk + 1; i < n; ++i) {
        const double f = lu[i * n + kk] / piv;
        if (f == 0.0) continue;
        lu[i * n + kk] = f;
        for (size_t j = kk + 1; j < n; ++j) lu[i * n + j] -= f * lu[kk * n + j];
      }
    }
    return true;
  }
  void project(const Lvl& v, Vec& a) const {
    if (v.nf == 0) return;
    double m = 0.0;
    for (size_t c = 0; c < a.size(); ++c) {
      if (v.fluid(c)) m += a[c];
      else a[c] = 0.0;
    }
    m /= double(v.nf);
    for (size_t c = 0; c < a.size(); ++c)
      if (v.fluid(c)) a[c] -= m;
  }
  void coarsest(const Vec& b, Vec& x) const {
    const Lvl& v = L.back();
    const size_t n = v.n;
    if (direct) {
      Vec y = b;
      project(v, y);
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < i; ++j) y[i] -= lu[i * n + j] * y[j];
      for (size_t i = n; i-- > 0;) {
        for (size_t j = i + 1; j < n; ++j) y[i] -= lu[i * n + j] * y[j];
        y[i] /= lu[i * n + i];
      }
      project(v, y);
      x = y;
      return;
    }
    x.assign(n, 0.0);
    for (int q = 0; q < 16; ++q) sweep(v, b, x);
  }
  static void sweep(const Lvl& v, const Vec& b, Vec& x) {
    const size_t n = x.size();
    Vec xn(n, 0.0);
    for (size_t i = 0; i < n; ++i)
      if (v.fluid(i)) xn[i] = x[i] + kOm * (b[i] - rowA(v, i, x)) / v.diag[i];
    x.swap(xn);
  }
  void cycle(size_t l, const Vec& b, Vec& x) const {
    const Lvl& v = L[l];
    if (l + 1 == L.size()) {
      coarsest(b, x);
      return;
    }
    x.assign(v.n, 0.0);
    sweep(v, b, x);
    sweep(v, b, x);
    const Lvl& c = L[l + 1];
    Vec bc(c.n, 0.0), ec;
    for (size_t i = 0; i < v.n; ++i)
      if (v.fluid(i)) bc[v.up[i]] += b[i] - rowA(v, i, x);
    cycle(l + 1, bc, ec);
    for (size_t i = 0; i < v.n; ++i) x[i] += ec[v.up[i]];
    sweep(v, b, x);
    sweep(v, b, x);
  }

  nbl_poisson_report solve(const Vec& bIn, Vec& xOut, double tol, int maxIt) const {
    const Lvl& v = L[0];
    const size_t n = v.n;
    Vec rhs(n), x(n, 0.0);
    for (size_t q = 0; q < n; ++q) rhs[q] = v.fluid(q) ? bIn[order[q]] : 0.0;
    project(v, rhs);
    const double bn = std::sqrt(dotv(rhs, rhs)) + 1e-300;
    nbl_poisson_report rep{0, 0.0, 0};
    Vec r = rhs;
    xOut.assign(n, 0.0);
    if (std::sqrt(dotv(r, r)) <= tol * bn) {
      rep.residual = std::sqrt(dotv(r, r)) / bn;
      rep.converged = 1;
      return rep;
    }
    auto minv = [&](const Vec& rr, Vec& zz) {
      cycle(0, rr, zz);
      project(v, zz);
    };
    Vec z, p, Ap(n), rn(n), zn;
    minv(r, z);
    p = z;
    double rz = dotv(r, z);
    int it = 0;
    for (; it < maxIt; ++it) {
      for (size_t i = 0; i < n; ++i) Ap[i] = v.fluid(i) ? rowA(v, i, p) : p[i];
      const double pAp = dotv(p, Ap);
      if (std::abs(pAp) < 1e-300) break;
      const double al = rz / pAp;
      for (size_t i = 0; i < n; ++i) {
        x[i] += al * p[i];
        rn[i] = r[i] - al * Ap[i];
      }
      project(v, rn);
      if (std::sqrt(dotv(rn, rn)) <= tol * bn) {
        ++it;
        r = rn;
        rep.converged = 1;
        break;
      }
      minv(rn, zn);
      double a = 0.0, c = 0.0;
      for (size_t i = 0; i < n; ++i) {
        a += zn[i] * rn[i];
        c += zn[i] * r[i];
      }
      const double be = (a - c) / rz;
      for (size_t i = 0; i < n; ++i) p[i] = zn[i] + be * p[i];
      rz = a;
      r = rn;
    }
    project(v, x);
    double rr = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const double ax = v.fluid(i) ? rowA(v, i, x) : x[i];
      const double e = rhs[i] - ax;
      rr += e * e;
    }
    rep.iterations = it;
    rep.residual = std::sqrt(rr) / bn;
    rep.converged = (rep.converged || rep.residual <= tol) ? 1 : 0;
    for (size_t q = 0; q < n; ++q) xOut[order[q]] = x[q];
    return rep;
  }
};

}  // namespace orc

using namespace orc;

extern "C" {

void* orc_ns_create(const nbl_grid_desc* d, const nbl_flow_cfg* cfg) {
  auto* s = new NS();
  s->g = makeGrid(d);
  s->cfg = *cfg;
  s->nu = cfg->velocity_scale * cfg->ref_length / cfg->reynolds;
  s->mg.tol = cfg->poisson_tol;
  s->mg.maxCyc = cfg->poisson_max_iters;
  for (uint8_t b : s->g.solid)
    if (b) {
      s->anySolid = true;
      break;
    }
  const Grid& g = s->g;
  s->u.assign(size_t(g.nx + 1) * g.ny * g.nz, 0.0);
  s->v.assign(size_t(g.nx) * (g.ny + 1) * g.nz, 0.0);
  s->w.assign(size_t(g.nx) * g.ny * (g.nz + 1), 0.0);
  s->p.assign(g.cells(), 0.0);
  return s;
}
void orc_ns_destroy(void* h) { delete static_cast<NS*>(h); }
void orc_ns_initialize(void* h) {  // ns_solver.cpp:95-112
  NS& s = *static_cast<NS*>(h);
  std::fill(s.u.begin(), s.u.end(), 0.0);
  std::fill(s.v.begin(), s.v.end(), 0.0);
  std::fill(s.w.begin(), s.w.end(), 0.0);
  std::fill(s.p.begin(), s.p.end(), 0.0);
  for (int f = 0; f < 6; ++f)
    if (s.g.face[f].type == NBL_BC_INLET && s.g.face[f].inlet_axis == 0) {
      std::fill(s.u.begin(), s.u.end(), s.g.face[f].inlet_speed);
      break;
    }
  s.bc(s.u, s.v, s.w, true);
  s.stepN = 0;
  s.t = 0.0;
  s.dtLast = 0.0;
}
int orc_ns_set_state(void* h, const double* u, size_t nu, const double* v, size_t nv, const double* w,
                     size_t nw, const double* p, size_t np, int step, double t, double dt_last) {
  NS& s = *static_cast<NS*>(h);
  if (nu != s.u.size() || nv != s.v.size() || nw != s.w.size() || np != s.p.size()) return -1;
  s.u.assign(u, u + nu);
  s.v.assign(v, v + nv);
  s.w.assign(w, w + nw);
  s.p.assign(p, p + np);
  s.stepN = step;
  s.t = t;
  s.dtLast = dt_last;
  return 0;
}
void orc_ns_get_state(void* h, double* u, double* v, double* w, double* p, int* step, double* t,
                      double* dt_last) {
  NS& s = *static_cast<NS*>(h);
  if (u) std::copy(s.u.begin(), s.u.end(), u);
  if (v) std::copy(s.v.begin(), s.v.end(), v);
  if (w) std::copy(s.w.begin(), s.w.end(), w);
  if (p) std::copy(s.p.begin(), s.p.end(), p);
  if (step) *step = s.stepN;
  if (t) *t = s.t;
  if (dt_last) *dt_last = s.dtLast;
}
double orc_ns_compute_dt(void* h) { return static_cast<NS*>(h)->cflDt(); }
void orc_ns_step(void* h, double dt, nbl_step_report* rep) {
  NS& s = *static_cast<NS*>(h);
  if (dt <= 0.0) dt = s.cflDt();
  const nbl_step_report r = s.advance(dt);
  if (rep) *rep = r;
}
void orc_ns_rhs(void* h, double* ru, double* rv, double* rw) {
  NS& s = *static_cast<NS*>(h);
  Vec a = s.u, b = s.v, c = s.w, x, y, z;
  s.bc(a, b, c, true);
  s.rhs(a, b, c, x, y, z);
  std::copy(x.begin(), x.end(), ru);
  std::copy(y.begin(), y.end(), rv);
  std::copy(z.begin(), z.end(), rw);
}
void orc_ns_cell_velocity(void* h, int comp, double* out) {
  NS& s = *static_cast<NS*>(h);
  for (int k = 0; k < s.g.nz; ++k)
    for (int j = 0; j < s.g.ny; ++j)
      for (int i = 0; i < s.g.nx; ++i)
        out[s.g.c(i, j, k)] = comp == 0 ? s.cU(i, j, k) : (comp == 1 ? s.cV(i, j, k) : s.cW(i, j, k));
}

// writeGridVtu (case.cpp:175-241): the ASCII snapshot as the reference prints
// it, through an ostream with precision 8.  Returns 0, or -1 when the file
// cannot be opened.
int orc_ns_write_vtu(void* h, const char* path) {
  NS& s = *static_cast<NS*>(h);
  const Grid& g = s.g;
  std::ofstream f(path);
  if (!f) return -1;
  f.precision(8);
  const size_t ncell = g.cells(), npts = ncell * 8;
  f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
       "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n  <UnstructuredGrid>\n";
  f << "    <Piece NumberOfPoints=\"" << npts << "\" NumberOfCells=\"" << ncell << "\">\n";
  f << "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const double x0 = g.org[0] + i * g.dx, x1 = x0 + g.dx;
        const double y0 = g.org[1] + j * g.dy, y1 = y0 + g.dy;
        const double z0 = g.org[2] + k * g.dz, z1 = z0 + g.dz;
        const double q[24] = {x0, y0, z0, x1, y0, z0, x1, y1, z0, x0, y1, z0,
                              x0, y0, z1, x1, y0, z1, x1, y1, z1, x0, y1, z1};
        for (int e = 0; e < 24; ++e) f << q[e] << (e == 23 ? '\n' : ' ');
      }
  f << "        </DataArray>\n      </Points>\n      <Cells>\n";
  f << "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
  for (size_t i = 0; i < npts; ++i) f << i << ((i + 1) % 8 == 0 ? '\n' : ' ');
  f << "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
  for (size_t i = 1; i <= ncell; ++i) f << i * 8 << ' ';
  f << "\n        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
  for (size_t i = 0; i < ncell; ++i) f << "12 ";
  f << "\n        </DataArray>\n      </Cells>\n      <CellData Scalars=\"speed\">\n";
  Vec cu(ncell), cv(ncell), cw(ncell), speed(ncell), solid(ncell);
  orc_ns_cell_velocity(h, 0, cu.data());
  orc_ns_cell_velocity(h, 1, cv.data());
  orc_ns_cell_velocity(h, 2, cw.data());
  for (size_t c = 0; c < ncell; ++c) {
    speed[c] = std::sqrt(cu[c] * cu[c] + cv[c] * cv[c] + cw[c] * cw[c]);
    solid[c] = g.solid[c] ? 1.0 : 0.0;
  }
  const std::pair<const char*, const Vec*> arrays[6] = {{"u", &cu},        {"v", &cv},         {"w", &cw},
                                                        {"p", &s.p},       {"speed", &speed}, {"solid", &solid}};
  for (const auto& [name, a] : arrays) {
    f << "        <DataArray type=\"Float64\" Name=\"" << name << "\" format=\"ascii\">\n";
    for (double x : *a) f << x << ' ';
    f << "\n        </DataArray>\n";
  }
  f << "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
  return f ? 0 : -1;
}

void orc_apply_a(const nbl_grid_desc* d, const double* x, double* out) {
  const Grid g = makeGrid(d);
  Vec xx(x, x + g.cells()), o;
  applyA(g, xx, o);
  std::copy(o.begin(), o.end(), out);
}
void orc_diagonal(const nbl_grid_desc* d, double* diag) {
  const Grid g = makeGrid(d);
  Vec dg;
  buildDiag(g, dg);
  std::copy(dg.begin(), dg.end(), diag);
}
void orc_mg_solve(const nbl_grid_desc* d, const double* b, double* x, double tol, int max_cycles,
                  nbl_poisson_report* rep) {
  const Grid g = makeGrid(d);
  Mg mg;
  mg.tol = tol;
  mg.maxCyc = max_cycles;
  Vec bb(b, b + g.cells()), xx;
  const nbl_poisson_report r = mg.solve(g, bb, xx);
  std::copy(xx.begin(), xx.end(), x);
  if (rep) *rep = r;
}
void orc_cg_solve(const nbl_grid_desc* d, const double* b, double* x, double tol, int max_iters,
                  nbl_poisson_report* rep) {
  const Grid g = makeGrid(d);
  Vec bb(b, b + g.cells()), xx;
  const nbl_poisson_report r = cgSolve(g, bb, xx, tol, max_iters);
  std::copy(xx.begin(), xx.end(), x);
  if (rep) *rep = r;
}
int orc_mg_levels(const nbl_grid_desc* d, int* dims) {
  const auto H = buildHier(makeGrid(d));
  if (dims)
    for (size_t l = 0; l < H->lv.size(); ++l) {
      dims[3 * l] = H->lv[l].g.nx;
      dims[3 * l + 1] = H->lv[l].g.ny;
      dims[3 * l + 2] = H->lv[l].g.nz;
    }
  return int(H->lv.size());
}
void orc_mg_smooth_once(const nbl_grid_desc* d, const double* b, double* x) {
  const auto H = buildHier(makeGrid(d));
  const size_t N = H->lv[0].g.cells();
  Vec bb(b, b + N), xx(x, x + N);
  jacobi(H->lv[0], bb, xx, 1, false);
  std::copy(xx.begin(), xx.end(), x);
}
void orc_mg_restrict(const nbl_grid_desc* d, const double* rf, double* rc) {
  const auto H = buildHier(makeGrid(d));
  if (H->lv.size() < 2) return;
  Vec r(rf, rf + H->lv[0].g.cells()), c;
  restrictTo(H->lv[0], H->lv[1], r, c, false, false);
  std::copy(c.begin(), c.end(), rc);
}
void orc_mg_prolong(const nbl_grid_desc* d, const double* ec, double* xf) {
  const auto H = buildHier(makeGrid(d));
  if (H->lv.size() < 2) return;
  Vec e(ec, ec + H->lv[1].g.cells()), x(xf, xf + H->lv[0].g.cells());
  prolongAdd(H->lv[1], H->lv[0], e, x);
  std::copy(x.begin(), x.end(), xf);
}

uint64_t orc_morton_encode(uint32_t x, uint32_t y, uint32_t z) { return menc(x, y, z); }
void orc_morton_decode(uint64_t key, uint32_t* x, uint32_t* y, uint32_t* z) {
  *x = gather(key);
  *y = gather(key >> 1);
  *z = gather(key >> 2);
}
void* orc_oct_create(const double origin[3], const double extent[3], int max_level) {
  try {
    return new Oct(origin, extent, max_level);
  } catch (...) {
    return nullptr;
  }
}
void orc_oct_destroy(void* h) { delete static_cast<Oct*>(h); }
size_t orc_oct_count(void* h) { return static_cast<Oct*>(h)->key.size(); }
void orc_oct_refine_uniform(void* h, int times) { static_cast<Oct*>(h)->refineUniform(times); }
void orc_oct_refine_box(void* h, const double lo[3], const double hi[3], int times) {
  static_cast<Oct*>(h)->refineBox(lo, hi, times);
}
int orc_oct_refine_key(void* h, uint64_t key) {
  try {
    static_cast<Oct*>(h)->refineKey(key);
    return 0;
  } catch (...) {
    return -1;
  }
}
int orc_oct_level_of(void* h, uint64_t key) {
  const Oct& o = *static_cast<Oct*>(h);
  const long i = o.find(key);
  return i < 0 ? -1 : int(o.lev[size_t(i)]);
}
void orc_oct_set_mask(void* h, size_t i, uint8_t m) { static_cast<Oct*>(h)->msk[i] = m; }
void orc_oct_dump(void* h, uint64_t* keys, uint8_t* levels, uint8_t* masks) {
  const Oct& o = *static_cast<Oct*>(h);
  for (size_t i = 0; i < o.key.size(); ++i) {
    if (keys) keys[i] = o.key[i];
    if (levels) levels[i] = o.lev[i];
    if (masks) masks[i] = o.msk[i];
  }
}
int orc_oct_face_neighbors(void* h, size_t i, int dir, uint64_t* out) {
  const auto v = static_cast<Oct*>(h)->nbrs(i, dir);
  for (size_t m = 0; m < v.size(); ++m) out[m] = v[m];
  return int(v.size());
}
void orc_oct_face_table(void* h, uint8_t* cnt, int64_t* nbr) {
  const Oct& o = *static_cast<Oct*>(h);
  for (size_t i = 0; i < o.key.size(); ++i)
    for (int d = 0; d < 6; ++d) {
      const auto v = o.nbrs(i, d);
      cnt[6 * i + d] = uint8_t(v.size());
      for (int m = 0; m < 4; ++m) nbr[24 * i + 4 * d + m] = m < int(v.size()) ? int64_t(v[m]) : -1;
    }
}
int orc_oct_is_balanced(void* h) { return static_cast<Oct*>(h)->balanced() ? 1 : 0; }
void orc_oct_centers(void* h, double* xyz) {
  const Oct& o = *static_cast<Oct*>(h);
  for (size_t i = 0; i < o.key.size(); ++i)
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] = o.center(i, a);
}
void orc_oct_apply(void* h, const double* x, double* out) {
  const Oct& o = *static_cast<Oct*>(h);
  Vec xx(x, x + o.key.size()), r;
  o.apply(xx, r);
  std::copy(r.begin(), r.end(), out);
}
void orc_oct_diagonal(void* h, double* diag) {
  const Oct& o = *static_cast<Oct*>(h);
  Vec d;
  o.diag(d);
  std::copy(d.begin(), d.end(), diag);
}
void orc_oct_solve(void* h, const double* b, double* x, double tol, int max_iters, nbl_poisson_report* rep) {
  const Oct& o = *static_cast<Oct*>(h);
  Vec bb(b, b + o.key.size()), xx;
  const nbl_poisson_report r = o.solve(bb, xx, tol, max_iters);
  std::copy(xx.begin(), xx.end(), x);
  if (rep) *rep = r;
}


void orc_oct_mg_solve(void* h, const double* b, double* x, double tol, int max_iters, nbl_poisson_report* rep) {
  const Oct& o = *static_cast<Oct*>(h);
  const OctMgOrc mg(o);
  Vec bb(b, b + o.key.size()), xx;
  const nbl_poisson_report r = mg.solve(bb, xx, tol, max_iters);
  std::copy(xx.begin(), xx.end(), x);
  if (rep) *rep = r;
}
int orc_oct_mg_levels(void* h, long long* counts, int cap) {
  const OctMgOrc mg(*static_cast<Oct*>(h));
  for (int l = 0; l < std::min(cap, int(mg.L.size())); ++l) counts[l] = (long long)mg.L[size_t(l)].n;
  return int(mg.L.size());
}


void orc_oct_sensors(void* h, const double* u, const double* v, const double* w, const double* p, double nu,
                     double* speedPrev, double* residPrev, double* out, uint8_t* flags) {
  const Oct& o = *static_cast<Oct*>(h);
  const size_t n = o.key.size();
  Vec U(u, u + n), V(v, v + n), W(w, w + n), P(p, p + n), sp(speedPrev, speedPrev + n), rp(residPrev, residPrev + n);
  o.sensors(U, V, W, P, nu, sp, rp, out, flags);
  std::copy(sp.begin(), sp.end(), speedPrev);
  std::copy(rp.begin(), rp.end(), residPrev);
}

double orc_oct_transport(void* h, const double* u, const double* v, const double* w, double* phi, double dt) {
  const Oct& o = *static_cast<Oct*>(h);
  const size_t n = o.key.size();
  Vec uu(u, u + n), vv(v, v + n), ww(w, w + n), pp(phi, phi + n);
  const double I = o.transport(uu, vv, ww, pp, dt);
  std::copy(pp.begin(), pp.end(), phi);
  return I;
}

}  // extern "C"
