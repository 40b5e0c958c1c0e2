// The adaptive control plane's flow seam — the reference's SolverBackend
// (core/include/nabla/adaptive/run.hpp:13-32) — and GpuBackend, the
// device-resident MAC solver behind it (in place of the reference's
// BaselineBackend, core/src/adaptive/run.cpp:23-65).
#pragma once

#include <cstddef>
#include <memory>

#include "nabla/flow/ns_solver.hpp"
#include "nabla/mesh/octree.hpp"

namespace nabla::adaptive {

using mesh::Vec3;

class SolverBackend {
 public:
  virtual ~SolverBackend() = default;
  virtual void step() = 0;
  virtual double sample(int component, Vec3 p) const = 0;  // 0=u,1=v,2=w,3=p
  virtual double cd() const = 0;
  virtual double cl() const = 0;
  virtual double momentumResidual() const = 0;
  virtual double nu() const = 0;
  virtual std::size_t uniformCells() const = 0;
  virtual double t() const = 0;
  virtual double dt() const = 0;
  virtual double cfl() const = 0;
  virtual double continuityResidual() const = 0;
  virtual double divergenceMax() const = 0;
  virtual double massError() const = 0;
  virtual int poissonIters() const = 0;
  virtual double poissonResidual() const = 0;
};

// The uniform MAC solve on the GPU: fields stay in HBM between steps; a
// sample() after a step copies the cell fields to the host once and reads
// the cell containing p (indices clamped into the grid, as the reference's
// baseline backend does).
class GpuBackend : public SolverBackend {
 public:
  GpuBackend(flow::UniformGrid grid, flow::FlowConfig cfg, flow::Precision precision = flow::Precision::FP64);
  void step() override;
  double sample(int component, Vec3 p) const override;
  double cd() const override { return last_.cd; }
  double cl() const override { return last_.cl; }
  double momentumResidual() const override { return last_.momentumResidual; }
  double nu() const override { return solver_->nu(); }
  std::size_t uniformCells() const override { return grid_.cellCount(); }
  double t() const override { return last_.t; }
  double dt() const override { return last_.dt; }
  double cfl() const override { return last_.cfl; }
  double continuityResidual() const override { return last_.continuityResidual; }
  double divergenceMax() const override { return last_.divergenceMax; }
  double massError() const override { return last_.massError; }
  int poissonIters() const override { return last_.poissonIters; }
  double poissonResidual() const override { return last_.poissonResidual; }
  [[nodiscard]] flow::NSSolver& solver() { return *solver_; }

 private:
  flow::UniformGrid grid_;
  std::unique_ptr<flow::NSSolver> solver_;
  flow::StepReport last_;
};

// Every leaf's u, v, w, p sampled from the backend at the leaf centre (the
// reference's sampleFlow, run.cpp:78-90), index-parallel to the tree.
struct LeafFields {
  std::vector<double> u, v, w, p;
};
LeafFields sampleFlow(const mesh::Octree& tree, const SolverBackend& backend);

}  // namespace nabla::adaptive
