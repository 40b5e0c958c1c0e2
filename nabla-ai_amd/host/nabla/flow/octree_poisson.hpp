// Cell-centred finite-volume pressure Poisson on octree leaves — the
// reference's OctreePoisson (core/include/nabla/flow/octree_poisson.hpp:29-55)
// over the device operator: (A p)_i = sum over fluid face neighbours j of
// (area / dist) (p_i - p_j), conservative across 2:1 jumps, identity rows on
// non-fluid leaves, a pure-Neumann (singular) system solved by Jacobi-PCG.
#pragma once

#include <string>
#include <vector>

#include "nabla/mesh/octree.hpp"

namespace nabla::flow {

class OctreePoisson {
 public:
  explicit OctreePoisson(double tol = 1e-9, int maxIters = 2000) : tol_(tol), maxIters_(maxIters) {}

  struct Report {
    int iterations = 0;
    double residual = 0.0;
    bool converged = false;
  };

  // Solve A x = b over the fluid leaves; x and b are parallel to the tree's
  // leaves, x is assigned.
  Report solve(const mesh::Octree& tree, const std::vector<double>& b, std::vector<double>& x);

  // Matrix-free operator action and its main diagonal (non-fluid leaves: 1).
  static void applyA(const mesh::Octree& tree, const std::vector<double>& x, std::vector<double>& out);
  static void buildDiagonal(const mesh::Octree& tree, std::vector<double>& diag);

  [[nodiscard]] std::string name() const { return "octree-cg-jacobi"; }

 private:
  double tol_;
  int maxIters_;
};

}  // namespace nabla::flow
