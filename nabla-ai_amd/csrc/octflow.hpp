// octflow.hpp — the flow step on octree leaves: cell-centred collocated
// finite volumes with a projection after each SSP-RK3 stage (see the
// nbl_octree_flow_* declarations in nbl_api.h for the scheme).
#pragma once

#include "octree.hpp"

namespace nbl {

class OctFlowDev {
 public:
  // The octree must outlive the flow.
  OctFlowDev(nbl_ctx* ctx, OctreeDev& tree, const nbl_octree_flow_cfg& cfg);
  ~OctFlowDev();
  OctFlowDev(const OctFlowDev&) = delete;
  OctFlowDev& operator=(const OctFlowDev&) = delete;
  void setState(const double* u, const double* v, const double* w, const double* p);
  void getState(double* u, double* v, double* w, double* p);
  nbl_octree_step_report step(double dt);
  void divergence(double* out);
  double cflDt();
  double cfl() const { return cfg_.cfl; }

 private:
  // Projection of the cell field in u_ with time step dt: face velocities,
  // pressure, corrected cell velocities.  Adds the solve's iterations to *it.
  void project(double dt, int* it, double* resid, double* ms);
  void faceInit();  // U_f = 0.5 (u_P + u_N).n (and boundary values)
  void fold(int K, unsigned maxMask, double* out);  // the partials of a reducing kernel

  nbl_ctx* ctx_;
  OctreeDev& tree_;
  nbl_octree_flow_cfg cfg_;
  hipStream_t s_;
  size_t n_;
  DevBuf u_[3], u0_[3], rhs_[3], p_, b_, U_, red_, scal_;
  DevBuf wl_;  // WENO-5 line (m2, m1, p1, p2) per first-neighbour face entry (WENO-5 runs)
  double* hscal_ = nullptr;  // pinned
  int step_ = 0;
  double t_ = 0.0;
  double inflow_ = 0.0, outflow_ = 0.0;
};

}  // namespace nbl

struct nbl_octree_flow {
  nbl_ctx* ctx;
  nbl_octree* tree;
  std::unique_ptr<nbl::OctFlowDev> impl;
};
