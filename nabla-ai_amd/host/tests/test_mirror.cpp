// The reference's own C++ tests, restated against this package's mirror
// headers (nabla/flow, nabla/mesh, nabla/adaptive) and run on the GPU:
//   core/tests/test_flow.cpp:21-61, 63-117 (CgPoisson statics and solve,
//   MgPoisson with an immersed solid), core/tests/test_octree_poisson.cpp:
//   47-145 (OctreePoisson on uniform and 2:1 trees, conservation, symmetry),
// plus the seams the mirror adds: NSSolver with a PoissonSolver outside the
// device path, and GpuBackend behind adaptive::SolverBackend.
// Exit status 0 when every check holds; one line per check.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "nabla/adaptive/run.hpp"
#include "nabla/flow/ns_solver.hpp"
#include "nabla/flow/octree_poisson.hpp"
#include "nabla/flow/poisson.hpp"
#include "nabla/mesh/octree.hpp"

using namespace nabla;
using namespace nabla::flow;
using nabla::mesh::Octree;
using nabla::mesh::Vec3;

namespace {
constexpr double kPi = 3.14159265358979323846;
int g_fail = 0;

void check(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  if (!ok) ++g_fail;
}

// test_flow.cpp:21-61
void cgManufacturedPeriodic() {
  UniformGrid g;
  g.nx = 16;
  g.ny = 16;
  g.nz = 1;
  g.dx = g.dy = g.dz = 1.0 / 16;
  for (auto& f : g.face) f.type = BoundaryType::Periodic;
  g.allocate();
  std::vector<double> phi(g.cellCount()), b;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      const double x = (i + 0.5) * g.dx, y = (j + 0.5) * g.dy;
      phi[g.cidx(i, j, 0)] = std::cos(2 * kPi * x) * std::cos(2 * kPi * y);
    }
  CgPoisson::applyA(g, phi, b);
  CgPoisson solver(1e-12, 5000);
  std::vector<double> x;
  const PoissonSolver::Report rep = solver.solve(g, b, x);
  check(rep.residual < 1e-8, "CgPoisson periodic: residual < 1e-8");
  double mp = 0, mx = 0;
  for (std::size_t c = 0; c < phi.size(); ++c) {
    mp += phi[c];
    mx += x[c];
  }
  mp /= double(phi.size());
  mx /= double(x.size());
  double err = 0;
  for (std::size_t c = 0; c < phi.size(); ++c) {
    const double e = (x[c] - mx) - (phi[c] - mp);
    err += e * e;
  }
  check(std::sqrt(err / double(phi.size())) < 1e-6, "CgPoisson periodic: error up to a constant < 1e-6");
  check(!CgPoisson::hasDirichlet(g), "CgPoisson::hasDirichlet: periodic box is singular");
}

UniformGrid blockChannel() {
  UniformGrid g;
  g.nx = 32;
  g.ny = 16;
  g.nz = 16;
  g.dx = g.dy = g.dz = 1.0 / 16;
  g.face[0].type = BoundaryType::Inlet;
  g.face[1].type = BoundaryType::Outlet;
  g.face[2].type = BoundaryType::Wall;
  g.face[3].type = BoundaryType::Wall;
  g.face[4].type = BoundaryType::Periodic;
  g.face[5].type = BoundaryType::Periodic;
  g.allocate();
  for (int k = 6; k < 10; ++k)
    for (int j = 0; j < 6; ++j)
      for (int i = 12; i < 18; ++i) g.solid[g.cidx(i, j, k)] = 1;
  return g;
}

// test_flow.cpp:63-117
void mgImmersedSolid() {
  const UniformGrid g = blockChannel();
  std::mt19937 rng(12345);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  std::vector<double> xExact(g.cellCount(), 0.0), b;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i)
        if (!g.isSolid(i, j, k)) xExact[g.cidx(i, j, k)] = dist(rng);
  CgPoisson::applyA(g, xExact, b);
  MgPoisson mg(1e-10, 100);
  std::vector<double> x;
  const PoissonSolver::Report rep = mg.solve(g, b, x);
  check(rep.converged, "MgPoisson immersed solid: converged");
  check(rep.residual < 1e-8, "MgPoisson immersed solid: residual < 1e-8");
  check(rep.iterations <= 50, "MgPoisson immersed solid: iterations <= 50 (" + std::to_string(rep.iterations) + ")");
  double err = 0.0;
  for (std::size_t c = 0; c < xExact.size(); ++c) {
    const double e = x[c] - xExact[c];
    err += e * e;
  }
  check(std::sqrt(err / double(xExact.size())) < 1e-6, "MgPoisson immersed solid: rms error < 1e-6");
  check(CgPoisson::hasDirichlet(g), "CgPoisson::hasDirichlet: the outlet makes it non-singular");
  std::vector<double> diag;
  CgPoisson::buildDiagonal(g, diag);
  const double h2 = 256.0;
  // an interior fluid cell couples to 6 fluid neighbours; a solid cell is an identity row
  check(diag[g.cidx(5, 8, 3)] == 6 * h2 && diag[g.cidx(13, 2, 7)] == 1.0, "CgPoisson::buildDiagonal: 6/h^2 and solid 1");
}

double manufactured(const Vec3& c) { return std::cos(kPi * c.x) * std::cos(kPi * c.y) * std::cos(kPi * c.z); }
double meanOver(const std::vector<double>& a) {
  double s = 0.0;
  for (double v : a) s += v;
  return s / double(a.size());
}
double l2UpToConstant(const std::vector<double>& a, const std::vector<double>& b) {
  const double ma = meanOver(a), mb = meanOver(b);
  double err = 0.0;
  for (std::size_t i = 0; i < a.size(); ++i) {
    const double e = (a[i] - ma) - (b[i] - mb);
    err += e * e;
  }
  return std::sqrt(err / double(a.size()));
}

// test_octree_poisson.cpp:47-65
void octreeUniform() {
  Octree tree(Vec3{0, 0, 0}, Vec3{1, 1, 1}, 4);
  tree.refineUniform(3);
  check(tree.cellCount() == 512, "Octree::refineUniform(3): 512 leaves");
  std::vector<double> pExact(tree.cellCount());
  for (std::size_t i = 0; i < tree.cellCount(); ++i) pExact[i] = manufactured(tree.cellCenter(i));
  std::vector<double> b;
  OctreePoisson::applyA(tree, pExact, b);
  OctreePoisson solver(1e-10, 5000);
  std::vector<double> x;
  const auto rep = solver.solve(tree, b, x);
  check(rep.residual < 1e-8, "OctreePoisson uniform: residual < 1e-8");
  check(l2UpToConstant(x, pExact) < 1e-6, "OctreePoisson uniform: error up to a constant < 1e-6");
}

void twoToOneTree(Octree& tree) {
  tree.refineUniform(2);
  tree.refineBox(Vec3{0, 0, 0}, Vec3{0.5, 0.5, 0.5}, 1);
}

// test_octree_poisson.cpp:67-95
void octreeLevelJump() {
  Octree tree(Vec3{0, 0, 0}, Vec3{1, 1, 1}, 5);
  twoToOneTree(tree);
  check(tree.isBalanced(), "Octree 2:1: balanced");
  int lo = 99, hi = 0;
  for (std::size_t i = 0; i < tree.cellCount(); ++i) {
    lo = std::min(lo, tree.level(i));
    hi = std::max(hi, tree.level(i));
  }
  check(hi > lo, "Octree 2:1: mixed levels");
  std::vector<double> pExact(tree.cellCount());
  for (std::size_t i = 0; i < tree.cellCount(); ++i) pExact[i] = manufactured(tree.cellCenter(i));
  std::vector<double> b;
  OctreePoisson::applyA(tree, pExact, b);
  OctreePoisson solver(1e-10, 5000);
  std::vector<double> x;
  const auto rep = solver.solve(tree, b, x);
  check(rep.residual < 1e-8, "OctreePoisson 2:1: residual < 1e-8");
  check(l2UpToConstant(x, pExact) < 1e-6, "OctreePoisson 2:1: error up to a constant < 1e-6");
}

// test_octree_poisson.cpp:97-145
void octreeConservativeSymmetric() {
  Octree tree(Vec3{0, 0, 0}, Vec3{1, 1, 1}, 5);
  twoToOneTree(tree);
  std::mt19937 rng(2024);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  std::vector<double> p(tree.cellCount());
  for (std::size_t i = 0; i < tree.cellCount(); ++i) p[i] = dist(rng);
  std::vector<double> Ap;
  OctreePoisson::applyA(tree, p, Ap);
  double sum = 0.0, maxAbs = 0.0;
  for (double v : Ap) {
    sum += v;
    maxAbs = std::max(maxAbs, std::abs(v));
  }
  check(std::abs(sum) < 1e-9 * (1.0 + maxAbs), "OctreePoisson: sum A p = 0 across level jumps");
  std::mt19937 rng2(7);
  std::vector<double> a(tree.cellCount()), bv(tree.cellCount());
  for (std::size_t i = 0; i < tree.cellCount(); ++i) {
    a[i] = dist(rng2);
    bv[i] = dist(rng2);
  }
  std::vector<double> Aa, Ab;
  OctreePoisson::applyA(tree, a, Aa);
  OctreePoisson::applyA(tree, bv, Ab);
  double aAb = 0.0, bAa = 0.0;
  for (std::size_t i = 0; i < a.size(); ++i) {
    aAb += a[i] * Ab[i];
    bAa += bv[i] * Aa[i];
  }
  check(std::abs(aAb - bAa) <= 1e-10 * std::abs(bAa), "OctreePoisson: symmetric");
  std::vector<double> diag;
  OctreePoisson::buildDiagonal(tree, diag);
  bool pos = true;
  for (double d : diag) pos = pos && d > 0.0;
  check(pos, "OctreePoisson::buildDiagonal: positive");
}

// A PoissonSolver outside the device path: it delegates to a device MG-PCG
// through host vectors and counts its calls.
class CountingPoisson : public PoissonSolver {
 public:
  explicit CountingPoisson(double tol) : inner_(tol, 5000) {}
  Report solve(const UniformGrid& g, const std::vector<double>& b, std::vector<double>& x) override {
    ++calls;
    return inner_.solve(g, b, x);
  }
  [[nodiscard]] std::string name() const override { return "counting"; }
  int calls = 0;

 private:
  MgPoisson inner_;
};

UniformGrid cubeChannel48() {
  UniformGrid g;
  g.nx = 48;
  g.ny = 16;
  g.nz = 16;
  g.dx = g.dy = g.dz = 0.25;
  g.face[0].type = BoundaryType::Inlet;
  g.face[0].inletSpeed = 1.0;
  g.face[1].type = BoundaryType::Outlet;
  g.face[2].type = g.face[3].type = BoundaryType::Wall;
  g.face[4].type = g.face[5].type = BoundaryType::Periodic;
  g.allocate();
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const V3 c = g.cellCenter(i, j, k);
        if (c.x >= 3 && c.x <= 4 && c.y >= 0 && c.y <= 1 && c.z >= 1.5 && c.z <= 2.5) g.solid[g.cidx(i, j, k)] = 1;
      }
  return g;
}

FlowConfig cubeConfig() {
  FlowConfig c;
  c.reynolds = 500.0;
  c.refArea = 1.0;
  c.cfl = 0.7;
  c.convection = ConvectionScheme::Weno5;
  c.poissonTol = 1e-10;
  return c;
}

// NSSolver takes any PoissonSolver (ns_solver.hpp:60-61): the same MG-PCG
// through the host seam gives the device path's fields bit for bit.
void nsSolverAnyPoisson() {
  const UniformGrid g = cubeChannel48();
  const FlowConfig cfg = cubeConfig();
  auto counting = std::make_shared<CountingPoisson>(cfg.poissonTol);
  NSSolver host(g, cfg, counting);
  NSSolver dev(g, cfg, std::make_shared<MgPoisson>(cfg.poissonTol, 5000));
  host.initialize();
  dev.initialize();
  StepReport rh, rd;
  for (int s = 0; s < 3; ++s) {
    rh = host.step();
    rd = dev.step();
  }
  check(counting->calls == 9, "NSSolver with a host PoissonSolver: three projections per step");
  check(rh.poissonIters == rd.poissonIters && rh.dt == rd.dt && rh.cd == rd.cd,
        "NSSolver with a host PoissonSolver: same step reports as the device path");
  check(host.uFace() == dev.uFace() && host.vFace() == dev.vFace() && host.wFace() == dev.wFace() &&
            host.pressure() == dev.pressure(),
        "NSSolver with a host PoissonSolver: fields bit-identical to the device path");
  NSSolver cg(g, cfg, std::make_shared<CgPoisson>(1e-12, 20000));
  cg.initialize();
  const StepReport rc = cg.step();
  check(rc.poissonIters > 0 && rc.continuityResidual < 1e-8, "NSSolver with CgPoisson: projected step");
}

void gpuBackend() {
  const UniformGrid g = cubeChannel48();
  adaptive::GpuBackend be(g, cubeConfig());
  adaptive::SolverBackend& sb = be;
  for (int s = 0; s < 3; ++s) sb.step();
  check(sb.poissonIters() > 0 && std::isfinite(sb.cd()) && sb.t() > 0.0 && sb.uniformCells() == g.cellCount(),
        "GpuBackend: steps and reports through SolverBackend");
  {
    // the host accessors against the device's cell-centred velocity, every cell
    const std::vector<double> cu = be.solver().cellVelocity(0), cv = be.solver().cellVelocity(1),
                              cw = be.solver().cellVelocity(2);
    bool same = true;
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          const std::size_t c = g.cidx(i, j, k);
          same = same && cu[c] == be.solver().cellU(i, j, k) && cv[c] == be.solver().cellV(i, j, k) &&
                 cw[c] == be.solver().cellW(i, j, k);
        }
    check(same, "NSSolver::cellU/V/W: the device's cell-centred velocity at every cell");
  }
  check(sb.sample(0, Vec3{0.6, 2.1, 3.3}) == be.solver().cellU(2, 8, 13) &&
            sb.sample(3, Vec3{100.0, -1.0, 0.1}) == be.solver().pressure()[g.cidx(47, 0, 0)],
        "GpuBackend::sample: the containing cell, indices clamped");
  Octree tree(Vec3{0, 0, 0}, Vec3{12, 4, 4}, 6);
  tree.refineUniform(3);
  tree.refineBox(Vec3{2, 0, 1}, Vec3{6, 2, 3}, 1);
  const adaptive::LeafFields f = adaptive::sampleFlow(tree, sb);
  bool same = f.u.size() == tree.cellCount();
  for (std::size_t i = 0; same && i < tree.cellCount(); ++i)
    same = f.w[i] == sb.sample(2, tree.cellCenter(i)) && f.p[i] == sb.sample(3, tree.cellCenter(i));
  check(same, "sampleFlow: every leaf sampled at its centre");
}

}  // namespace

int main() {
  const std::pair<const char*, void (*)()> cases[] = {
      {"cg periodic", cgManufacturedPeriodic},         {"mg immersed solid", mgImmersedSolid},
      {"octree uniform", octreeUniform},               {"octree 2:1", octreeLevelJump},
      {"octree conservation", octreeConservativeSymmetric}, {"NSSolver any PoissonSolver", nsSolverAnyPoisson},
      {"GpuBackend", gpuBackend}};
  for (const auto& [name, fn] : cases) {
    try {
      fn();
    } catch (const std::exception& e) {
      check(false, std::string(name) + ": " + e.what());
    }
  }
  std::printf("%d failed\n", g_fail);
  return g_fail == 0 ? 0 : 1;
}
