// Pressure-Poisson plugin seam — the reference's PoissonSolver interface
// (core/include/nabla/flow/poisson.hpp:20-32) with device implementations.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "nabla/flow/grid.hpp"

namespace nabla::flow {

class PoissonSolver {
 public:
  virtual ~PoissonSolver() = default;
  struct Report {
    int iterations = 0;
    double residual = 0.0;
    bool converged = false;
  };
  // x and b are compact interior arrays of length grid.cellCount(); x is assigned.
  virtual Report solve(const UniformGrid& grid, const std::vector<double>& b, std::vector<double>& x) = 0;
  [[nodiscard]] virtual std::string name() const = 0;
};

enum class Precision { FP64 = NBL_PREC_FP64, Mixed = NBL_PREC_MIXED, MixedBF16 = NBL_PREC_MIXED_BF16 };

// Device solver behind the plugin seam; rebuilt only when the grid changes
// (dims, boundary kinds, solid mask), like MgPoisson's cache (multigrid.cpp:208-266).
class DevicePoisson : public PoissonSolver {
 public:
  DevicePoisson(int kind, double tol, int maxIters, int pre, int post, Precision prec);
  ~DevicePoisson() override;
  Report solve(const UniformGrid& grid, const std::vector<double>& b, std::vector<double>& x) override;
  [[nodiscard]] double tol() const { return tol_; }
  [[nodiscard]] int maxIters() const { return maxIters_; }
  [[nodiscard]] nbl_solver_opts opts() const;

 private:
  int kind_;
  double tol_;
  int maxIters_;
  int pre_, post_;
  Precision prec_;
  nbl_poisson* ps_ = nullptr;
  std::uint64_t sig_ = 0;
};

// MG-PCG: flexible CG preconditioned by one V(pre,post) cycle ("mg-pcg").
class HipMgPoisson : public DevicePoisson {
 public:
  explicit HipMgPoisson(double tol = 1e-9, int maxCycles = 100, int preSmooth = 2, int postSmooth = 2,
                        Precision prec = Precision::FP64)
      : DevicePoisson(0, tol, maxCycles, preSmooth, postSmooth, prec) {}
  [[nodiscard]] std::string name() const override { return "mg-pcg"; }
};

// Jacobi-preconditioned CG ("cg-jacobi").
class HipCgPoisson : public DevicePoisson {
 public:
  explicit HipCgPoisson(double tol = 1e-10, int maxIters = 5000)
      : DevicePoisson(1, tol, maxIters, 2, 2, Precision::FP64) {}
  [[nodiscard]] std::string name() const override { return "cg-jacobi"; }

  // CgPoisson's statics (poisson.hpp:47-52): A x for the discrete -Laplacian
  // with the grid's closure (identity rows on solid cells), whether an outlet
  // makes the system non-singular, and A's main diagonal (solid cells 1).
  // The operator and diagonal are evaluated on the device, bit-identical to
  // the reference's (poisson.cpp:78-134); out and diag are assigned.
  static void applyA(const UniformGrid& grid, const std::vector<double>& x, std::vector<double>& out);
  static bool hasDirichlet(const UniformGrid& grid);
  static void buildDiagonal(const UniformGrid& grid, std::vector<double>& diag);
};

// The reference's class names for the same solvers.
using MgPoisson = HipMgPoisson;
using CgPoisson = HipCgPoisson;

}  // namespace nabla::flow
