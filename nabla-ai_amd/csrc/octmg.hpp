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
  // Arithmetic of the cycle: NBL_PREC_FP64 (level vectors and operators in
  // FP64), NBL_PREC_MIXED (FP32 copies of both) or NBL_PREC_MIXED_BF16 (FP32,
  // with level 0's pre-smoothed iterate stored as BF16 between the sweeps
  // that write and read it).  FP32 copies are made on first use.
  void setPrecision(int prec, hipStream_t s);
  int precision() const { return prec_; }
  bool fp32() const { return prec_ != 0; }
  // Level 0's right-hand side (FP32 forms: the caller writes float(r) here;
  // the FP64 form reads the caller's r) and its first sweep from a zero
  // guess: the caller writes omega*r/diag (fluid leaves, 0 elsewhere) in the
  // cycle's type before vcycle().  diag0() is level 0's diagonal in that type.
  float* rhs0F() { return lv_[0].bF.as<float>(); }
  void* firstSweep() { return fp32() ? lv_[0].xaF.as<void>() : lv_[0].xa.as<void>(); }
  const void* diag0() const;
  // One V(2,2) cycle for right-hand side r on level 0 (FP64 form; the FP32
  // forms read rhs0F()).  Its result (double or float as the form) is
  // result(); the last kernel writes `partials()` per-block partial fluid
  // sums of it into rb.
  void vcycle(const double* r, RedBuf rb, const SolveCtl* ctl, hipStream_t s);
  // The same cycle whose last level-0 sweep writes partials of result.r,
  // result.rOld and the result's fluid sum (rows 0, 1, 2) instead: the
  // flexible CG's dots with the result's mean left to the caller.
  void vcycleDots(const double* r, const double* rOld, RedBuf rb, const SolveCtl* ctl, hipStream_t s);
  const void* result() const { return res0_; }
  int partials() const { return parts_; }

 private:
  struct Lvl {
    OctLevel own;               // coarse levels
    const OctLevel* op = nullptr;
    DevBuf up;                  // i32[n]: parent in level l+1
    DevBuf cstart;              // i32[n+1]: children of cell I at child[cstart[I] .. cstart[I+1])
    DevBuf child;               // i32[n of level l-1]: cells of level l-1 grouped by parent
    DevBuf b, xa, xb;           // FP64 level vectors
    DevBuf scoefF, xcoefF, diagF, bF, xaF, xbF;  // FP32 operator copy and level vectors
    DevBuf ctabF;               // FP32 class values (levels with coefficient classes)
    DevBuf xbB;                 // level 0, BF16 form: the pre-smoothed iterate
    int S[3] = {0, 0, 0};       // target log2 box size per axis of this level
  };
  void buildCoarse(const OctreeDev& fine, hipStream_t s);
  bool buildDirect(hipStream_t s);
  template <class T>
  OctViewT<T> view(size_t l) const;
  template <class T>
  void cycle(size_t l, const T* b, RedBuf rb, const SolveCtl* ctl, hipStream_t s, const T** out);

  nbl_ctx* ctx_;
  std::vector<Lvl> lv_;
  bool direct_ = false;
  int prec_ = 0;
  DevBuf code_, M_, Mf_;
  const void* res0_ = nullptr;
  const double* dRn_ = nullptr;  // vcycleDots' r and rOld while its cycle is enqueued
  const double* dR_ = nullptr;
  int parts_ = 0;
  double buildMs_ = 0.0;
};

}  // namespace nbl
