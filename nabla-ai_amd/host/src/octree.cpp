// Host layer over the octree entry points: the Octree leaf set
// (nbl_tree_*), OctreePoisson (nbl_octree_*) and the adaptive control plane's
// GPU backend.
#include <algorithm>
#include <stdexcept>

#include "nabla/adaptive/run.hpp"
#include "nabla/flow/device.hpp"
#include "nabla/flow/octree_poisson.hpp"
#include "nabla/flow/poisson.hpp"
#include "nabla/mesh/octree.hpp"

namespace nabla::mesh {

Octree::Octree(Vec3 origin, Vec3 extent, int maxLevel) : origin_(origin), extent_(extent), maxLevel_(maxLevel) {
  const double o[3] = {origin.x, origin.y, origin.z}, e[3] = {extent.x, extent.y, extent.z};
  flow::nblCheck(nbl_tree_create(o, e, maxLevel, &tree_), "Octree");
  refresh();
}

Octree::~Octree() {
  if (tree_) nbl_tree_destroy(tree_);
}

void Octree::refresh() {
  size_t n = 0;
  flow::nblCheck(nbl_tree_count(tree_, &n), "Octree");
  key_.resize(n);
  level_.resize(n);
  centre_.resize(3 * n);
  flow::nblCheck(nbl_tree_leaves(tree_, key_.data(), level_.data(), centre_.data()), "Octree");
  mask_.assign(n, static_cast<uint8_t>(CellMask::Fluid));
}

void Octree::refineUniform(int times) {
  flow::nblCheck(nbl_tree_refine_uniform(tree_, times), "Octree::refineUniform");
  refresh();
}

void Octree::refineBox(Vec3 lo, Vec3 hi, int times) {
  for (int t = 0; t < times; ++t) {
    std::vector<uint8_t> f(level_.size(), 0);
    bool any = false;
    for (std::size_t i = 0; i < f.size(); ++i) {
      const Vec3 c = cellCenter(i);
      if (c.x >= lo.x && c.x <= hi.x && c.y >= lo.y && c.y <= hi.y && c.z >= lo.z && c.z <= hi.z &&
          level_[i] < maxLevel_) {
        f[i] = 1;
        any = true;
      }
    }
    if (!any) return;
    flow::nblCheck(nbl_tree_refine_flagged(tree_, f.data(), f.size()), "Octree::refineBox");
    refresh();
  }
}

bool Octree::isBalanced() const {
  int b = 0;
  flow::nblCheck(nbl_tree_is_balanced(tree_, &b), "Octree::isBalanced");
  return b != 0;
}

}  // namespace nabla::mesh

namespace nabla::flow {
namespace {
// A device octree for one call: leaves, masks and the domain of the tree.
struct DevTree {
  nbl_octree* oc = nullptr;
  explicit DevTree(const mesh::Octree& t) {
    const mesh::Vec3 o = t.origin(), e = t.extent();
    const double od[3] = {o.x, o.y, o.z}, ed[3] = {e.x, e.y, e.z};
    nblCheck(nbl_octree_create(deviceContext(), t.keys().data(), t.levels().data(), t.masks().data(), t.cellCount(),
                               od, ed, t.maxLevel(), &oc),
             "OctreePoisson");
  }
  ~DevTree() {
    if (oc) nbl_octree_destroy(oc);
  }
};
}  // namespace

OctreePoisson::Report OctreePoisson::solve(const mesh::Octree& tree, const std::vector<double>& b,
                                           std::vector<double>& x) {
  if (b.size() != tree.cellCount()) throw std::invalid_argument("OctreePoisson::solve: b size mismatch");
  DevTree d(tree);
  x.assign(tree.cellCount(), 0.0);
  nbl_poisson_report r{};
  nblCheck(nbl_octree_poisson_solve(d.oc, b.data(), x.data(), tol_, maxIters_, &r), "OctreePoisson::solve");
  return Report{r.iterations, r.residual, r.converged != 0};
}

void OctreePoisson::applyA(const mesh::Octree& tree, const std::vector<double>& x, std::vector<double>& out) {
  if (x.size() != tree.cellCount()) throw std::invalid_argument("OctreePoisson::applyA: x size mismatch");
  DevTree d(tree);
  out.assign(tree.cellCount(), 0.0);
  nblCheck(nbl_octree_apply(d.oc, x.data(), out.data()), "OctreePoisson::applyA");
}

void OctreePoisson::buildDiagonal(const mesh::Octree& tree, std::vector<double>& diag) {
  DevTree d(tree);
  diag.assign(tree.cellCount(), 1.0);
  nblCheck(nbl_octree_diagonal(d.oc, diag.data()), "OctreePoisson::buildDiagonal");
}

}  // namespace nabla::flow

namespace nabla::adaptive {

GpuBackend::GpuBackend(flow::UniformGrid grid, flow::FlowConfig cfg, flow::Precision precision)
    : grid_(std::move(grid)) {
  if (grid_.solid.empty()) grid_.allocate();
  auto poisson = std::make_shared<flow::HipMgPoisson>(cfg.poissonTol, cfg.poissonMaxIters, 2, 2, precision);
  solver_ = std::make_unique<flow::NSSolver>(grid_, cfg, poisson);
  solver_->initialize();
}

void GpuBackend::step() { last_ = solver_->step(); }

double GpuBackend::sample(int component, Vec3 p) const {  // run.cpp:45-58
  int i = static_cast<int>((p.x - grid_.origin.x) / grid_.dx);
  int j = static_cast<int>((p.y - grid_.origin.y) / grid_.dy);
  int k = static_cast<int>((p.z - grid_.origin.z) / grid_.dz);
  i = std::clamp(i, 0, grid_.nx - 1);
  j = std::clamp(j, 0, grid_.ny - 1);
  k = std::clamp(k, 0, grid_.nz - 1);
  switch (component) {
    case 0: return solver_->cellU(i, j, k);
    case 1: return solver_->cellV(i, j, k);
    case 2: return solver_->cellW(i, j, k);
    default: return solver_->pressure()[grid_.cidx(i, j, k)];
  }
}

LeafFields sampleFlow(const mesh::Octree& tree, const SolverBackend& backend) {  // run.cpp:78-90
  LeafFields f;
  const std::size_t n = tree.cellCount();
  f.u.resize(n);
  f.v.resize(n);
  f.w.resize(n);
  f.p.resize(n);
  for (std::size_t i = 0; i < n; ++i) {
    const Vec3 c = tree.cellCenter(i);
    f.u[i] = backend.sample(0, c);
    f.v[i] = backend.sample(1, c);
    f.w[i] = backend.sample(2, c);
    f.p[i] = backend.sample(3, c);
  }
  return f;
}

}  // namespace nabla::adaptive
