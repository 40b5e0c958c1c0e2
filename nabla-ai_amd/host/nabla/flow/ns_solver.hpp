// MAC Navier-Stokes solver with the reference's public surface
// (core/include/nabla/flow/ns_solver.hpp:12-156); state is resident in HBM and
// copied to the host only for the accessors, checkpoints and snapshots.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "nabla/flow/grid.hpp"
#include "nabla/flow/poisson.hpp"

namespace nabla::flow {

enum class ConvectionScheme { Central = NBL_CONV_CENTRAL, Weno5 = NBL_CONV_WENO5 };

struct FlowConfig {
  double reynolds = 100.0;
  double refLength = 1.0;
  double velocityScale = 1.0;
  double density = 1.0;
  double cfl = 0.8;
  V3 bodyForce{};
  ConvectionScheme convection = ConvectionScheme::Central;
  double poissonTol = 1e-9;
  int poissonMaxIters = 5000;
  double tolContinuity = 1e-7;
  double tolMass = 1e-8;
  double refArea = 0.0;
};

struct StepReport {
  int step = 0;
  double t = 0.0;
  double dt = 0.0;
  double cfl = 0.0;
  double momentumResidual = 0.0;
  double continuityResidual = 0.0;
  double divergenceMax = 0.0;
  double massError = 0.0;
  int poissonIters = 0;
  double poissonResidual = 0.0;
  double cd = 0.0;
  double cl = 0.0;
  std::size_t cellCount = 0;
  bool accepted = false;
};

class NSSolver {
 public:
  // `poisson` may be null (MG-PCG with the config's tolerance), a device
  // MG-PCG (HipMgPoisson), whose tolerance, cycle limit and precision the
  // device pressure solve uses with the state kept in HBM, or any other
  // PoissonSolver (ns_solver.hpp:60-61 of the reference takes any): each
  // projection then copies its right-hand side to the host, calls
  // poisson->solve(grid, b, x) and copies x back.
  NSSolver(UniformGrid grid, FlowConfig config, std::shared_ptr<PoissonSolver> poisson = nullptr);
  ~NSSolver();
  NSSolver(const NSSolver&) = delete;
  NSSolver& operator=(const NSSolver&) = delete;

  void initialize();
  StepReport step();
  StepReport step(double dt);

  [[nodiscard]] double time() const { return t_; }
  [[nodiscard]] int stepIndex() const { return step_; }
  [[nodiscard]] double dtLast() const { return dtLast_; }
  [[nodiscard]] double nu() const { return nu_; }
  [[nodiscard]] const UniformGrid& grid() const { return grid_; }
  [[nodiscard]] const FlowConfig& config() const { return cfg_; }

  [[nodiscard]] std::vector<double> cellVelocity(int component) const;
  [[nodiscard]] std::vector<double> cellPressure() const { return pressure(); }
  [[nodiscard]] double cellU(int i, int j, int k) const;
  [[nodiscard]] double cellV(int i, int j, int k) const;
  [[nodiscard]] double cellW(int i, int j, int k) const;

  [[nodiscard]] const std::vector<double>& uFace() const;
  [[nodiscard]] const std::vector<double>& vFace() const;
  [[nodiscard]] const std::vector<double>& wFace() const;
  [[nodiscard]] const std::vector<double>& pressure() const;
  void setState(std::vector<double> u, std::vector<double> v, std::vector<double> w, std::vector<double> p,
                int step, double t, double dtLast);

  [[nodiscard]] double computeDt() const;

 private:
  void fetch() const;

  static int hostSolve(void* self, const double* b, double* x, std::size_t n, nbl_poisson_report* rep);

  UniformGrid grid_;
  FlowConfig cfg_;
  double nu_;
  std::shared_ptr<PoissonSolver> hostPoisson_;  // a solver outside the device path
  nbl_run* run_ = nullptr;
  int step_ = 0;
  double t_ = 0.0;
  double dtLast_ = 0.0;
  // host mirror, refreshed on demand after a step
  mutable bool fresh_ = false;
  mutable std::vector<double> u_, v_, w_, p_;
};

}  // namespace nabla::flow
