// octmg.hpp — geometric multigrid on octree leaves.
//
// Level 0 is the tree itself.  Each coarser level merges cells into boxes of a
// target size that grows first along the axes where the cells are physically
// smallest (semi-coarsening for anisotropic boxes, plain octree coarsening
// for cubes); the coarse operator is the two-point finite-volume one (summed
// face areas / centre distance), a coarse cell is fluid iff any child is
// fluid, residuals restrict by summation and corrections prolong by
// injection.  On a uniform all-fluid tree in a cube this is exactly the
// uniform solver's hierarchy with wall boundaries: cell-volume scaling
// cancels between the restricted residual and the coarse operator.
#pragma once

#include <vector>

#include "octree.hpp"

namespace nbl {

class OctMg {
 public:
  OctMg(nbl_ctx* ctx, const OctreeDev& fine, hipStream_t s);
  int levels() const { return int(lv_.size()); }
  long long count(int l) const { return (long long)lv_[size_t(l)].op->n; }
  double buildMs() const { return buildMs_; }
  bool directCoarse() const { return direct_; }
  // First sweep of level 0 from a zero guess: the caller writes omega*r/diag
  // (fluid leaves, 0 elsewhere) here before vcycle().
  double* firstSweep() { return lv_[0].xa.as<double>(); }
  // One V(2,2) cycle for right-hand side r on level 0.  Its result is
  // result(); the last kernel writes `partials()` per-block partial fluid sums
  // of it into rb.
  void vcycle(const double* r, RedBuf rb, const SolveCtl* ctl, hipStream_t s);
  const double* result() const { return res0_; }
  int partials() const { return parts_; }

 private:
  struct Lvl {
    OctLevel own;               // coarse levels
    const OctLevel* op = nullptr;
    DevBuf up;                  // i32[n]: parent in level l+1
    DevBuf cstart;              // i32[n+1]: children of cell I at child[cstart[I] .. cstart[I+1])
    DevBuf child;               // i32[n of level l-1]: cells of level l-1 grouped by parent
    DevBuf b, xa, xb;
    int S[3] = {0, 0, 0};       // target log2 box size per axis of this level
  };
  void buildCoarse(const OctreeDev& fine, hipStream_t s);
  bool buildDirect(hipStream_t s);
  void cycle(size_t l, const double* b, RedBuf rb, const SolveCtl* ctl, hipStream_t s, const double** out);

  nbl_ctx* ctx_;
  std::vector<Lvl> lv_;
  bool direct_ = false;
  DevBuf code_, M_, Mf_;
  const double* res0_ = nullptr;
  int parts_ = 0;
  double buildMs_ = 0.0;
};

}  // namespace nbl
