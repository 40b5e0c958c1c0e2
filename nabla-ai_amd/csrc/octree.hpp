// octree.hpp — device octree: Morton-sorted leaves, face tables, CSR operator
// and the Jacobi-PCG pressure solve on the leaves.
#pragma once

#include "solver.hpp"

namespace nbl {

class OctreeDev {
 public:
  OctreeDev(nbl_ctx* ctx, const uint64_t* keys, const uint8_t* levels, const uint8_t* masks, size_t n,
            const double origin[3], const double extent[3], int maxLevel);
  ~OctreeDev();
  size_t size() const { return n_; }
  size_t nnz() const { return nnz_; }
  void permutation(uint64_t* perm);
  void faceTable(uint8_t* count, int32_t* nbr);
  // caller-order host arrays
  void apply(const double* x, double* out);
  void diagonal(double* out);
  nbl_poisson_report solve(const double* b, double* x, double tol, int maxIters);
  // sorted-order device arrays
  void applySorted(const double* x, double* out, hipStream_t s);
  nbl_poisson_report solveSorted(const double* b, double* x, double tol, int maxIters, hipStream_t s);
  hipStream_t stream() const { return stream_; }
  // Device time of construction (sort, validation, face table, CSR) and of the
  // last solve (ms), and a timing loop over one kernel (0: the SpMV).
  double setupMs() const { return setupMs_; }
  double lastSolveMs() const { return lastMs_; }
  void profile(int kernel, int reps, double* ms, double* bytes);

 private:
  void toSorted(const double* hostIn, double* dSorted);
  void fromSorted(const double* dSorted, double* hostOut);

  nbl_ctx* ctx_;
  size_t n_ = 0, nnz_ = 0, nf_ = 0;
  int maxL_ = 0;
  double org_[3], fine_[3];
  hipStream_t stream_ = nullptr;
  DevBuf keys_, perm_, lev_, mask_, cnt_, nbr_, rowPtr_, col_, coef_, diag_;
  DevBuf x_, rhs_, r_, z_, p_, Ap_, t0_, t1_;
  DevBuf ctl_, red_;
  SolveCtl* hctl_ = nullptr;
  int predicted_ = 16;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  double setupMs_ = 0.0, lastMs_ = 0.0;
};

}  // namespace nbl

struct nbl_octree {
  nbl_ctx* ctx;
  std::unique_ptr<nbl::OctreeDev> impl;
};
