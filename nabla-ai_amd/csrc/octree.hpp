// octree.hpp — device octree: Morton-sorted leaves, face tables, CSR operator,
// the Jacobi-PCG pressure solve on the leaves and the multigrid-preconditioned
// flexible CG whose levels are the tree's own coarsenings.
#pragma once

#include <memory>

#include "solver.hpp"

namespace nbl {

// One set of Morton-sorted leaves with its fluid-fluid operator in CSR form:
// row i holds the fluid neighbours of fluid leaf i in the reference's
// accumulation order (direction -x,+x,-y,+y,-z,+z, then face order) and the
// face coefficients area/dist; diag is their sum (1 on non-fluid rows and on
// fluid rows without fluid neighbours).
struct OctLevel {
  size_t n = 0, nnz = 0, nf = 0, nx = 0;
  DevBuf keys, lev, mask, rowPtr, col, coef, diag;
  // Cell boxes in finest-index units: keys = Morton code of the anchor, lsz =
  // log2 size per axis (x | y << 8 | z << 16; octree leaves are cubes).  Per
  // coupling: the shared face area and the axis (u8).
  DevBuf lsz, area, cax;
  // The same couplings for the operator kernels: the first kOctSlots of every
  // row in slot-major arrays (slot s of row i at s*n + i, so a wavefront's
  // reads of one slot coalesce), the rest in an overflow CSR.
  DevBuf cnt, scol, scoef, xrow, xcol, xcoef;
  // Coefficient classes: where a level's coefficients take at most
  // kOctClasses distinct values (the usual case: face areas and centre
  // distances of dyadic boxes), ctab holds them (ascending by bit pattern)
  // and meta holds 8 bytes per leaf: the class of slot s in byte s (s < 6)
  // and the row's coupling count in byte 6.  ncls == 0: no classes, the
  // kernels read scoef.
  DevBuf meta, ctab;
  int ncls = 0;
};

constexpr int kOctClasses = 256;

constexpr int kOctSlots = 6;
// Leaves an octree may hold: its face table (24 entries per leaf) and its
// coupling CSR (at most 24 per leaf) are indexed in 32 bits.
constexpr long long kMaxOctreeLeaves = ((1LL << 31) - 1) / 24;

// Morton code with 21 bits per axis: x at bit 3i, y at 3i+1, z at 3i+2
// (morton.hpp:15-48).
__host__ __device__ inline uint64_t mortonSpread(uint32_t a) {
  uint64_t x = a & 0x1fffffull;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__host__ __device__ inline uint32_t mortonAxis(uint64_t key, int axis) {
  uint64_t x = (key >> axis) & 0x1249249249249249ull;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
  x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
  x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
  x = (x ^ (x >> 32)) & 0x1fffffull;
  return uint32_t(x);
}
__host__ __device__ inline uint64_t mortonEnc3(uint32_t x, uint32_t y, uint32_t z) {
  return mortonSpread(x) | (mortonSpread(y) << 1) | (mortonSpread(z) << 2);
}

// Geometry of a tree's index space: origin, finest cell size per axis and the
// finest level.
struct Geo {
  double org[3], fine[3];
  int maxL;
};
// Reference centre (octree.cpp:84-91) and size (octree.cpp:97-100) of a leaf.
__device__ __forceinline__ double centerAxis(const Geo& G, uint64_t key, int lvl, int axis) {
  const uint32_t a = mortonAxis(key, axis);
  const uint32_t sz = 1u << (G.maxL - lvl);
  const double units = double(a) + 0.5 * double(sz);
  return G.org[axis] + units * G.fine[axis];
}
__device__ __forceinline__ double sizeAxis(const Geo& G, int lvl, int axis) {
  return double(1u << (G.maxL - lvl)) * G.fine[axis];
}

// The face table as the per-leaf kernels walk it: face counts per direction
// (u8[6n], directions -x,+x,-y,+y,-z,+z), the first neighbour of each
// direction slot-major (first[d*n + i], -1 when none), and the further
// neighbours of leaf i in direction order from xnb[xrow[i]].
struct FaceWalk {
  const uint8_t* __restrict__ cnt;
  const int32_t* __restrict__ first;
  const int32_t* __restrict__ xrow;
  const int32_t* __restrict__ xnb;
};

// Device view of a level's operator with coefficients and diagonal in T
// (double: the operator itself; float: its FP32 copy for the mixed-precision
// V-cycle).
template <class T>
struct OctViewT {
  long long n;
  const uint8_t* mask;
  const uint8_t* cnt;
  const int32_t* scol;
  const T* scoef;
  const int32_t* xrow;
  const int32_t* xcol;
  const T* xcoef;
  const T* diag;
  const uint2* meta;  // coefficient classes and counts (null: read cnt and scoef)
  const T* ctab;      // class values
};
using OctView = OctViewT<double>;
inline OctView octView(const OctLevel& L) {
  return OctView{(long long)L.n,       L.mask.as<uint8_t>(),  L.cnt.as<uint8_t>(),  L.scol.as<int32_t>(),
                 L.scoef.as<double>(), L.xrow.as<int32_t>(),  L.xcol.as<int32_t>(), L.xcoef.as<double>(),
                 L.diag.as<double>(),  L.ncls ? L.meta.as<uint2>() : nullptr,  L.ctab.as<double>()};
}

// sum_e coef_e (x_i - x_col_e) over row i's couplings in the reference's order
// (octree_poisson.cpp:64-85), x read through X(j), in T arithmetic.
// With coefficient classes one 8-byte load gives the count and every slot's
// class, and the coefficient is the class value (the same bits as scoef).
// Every slot holds a valid column (an unused slot names row i itself, with a
// zero coefficient), so all six columns, coefficients and neighbour values are
// loaded before the count is known: two memory round trips per row (columns,
// then neighbour values) instead of one per dependent step.  Only the first
// count slots are accumulated, in slot order.
template <class T, class Rd>
__device__ __forceinline__ T octRowA(const OctViewT<T>& L, long long i, const Rd& X) {
  const T xi = X(i);
  T acc = T(0);
  int c;
  int32_t col[kOctSlots];
  T xv[kOctSlots], cf[kOctSlots];
#pragma unroll
  for (int s = 0; s < kOctSlots; ++s) col[s] = L.scol[(long long)s * L.n + i];
  if (L.meta) {
    const uint2 m = L.meta[i];
#pragma unroll
    for (int s = 0; s < kOctSlots; ++s) xv[s] = X(col[s]);
    c = int((m.y >> 16) & 0xffu);
#pragma unroll
    for (int s = 0; s < kOctSlots; ++s) cf[s] = L.ctab[((s < 4 ? m.x : m.y) >> (8 * (s & 3))) & 0xffu];
  } else {
#pragma unroll
    for (int s = 0; s < kOctSlots; ++s) cf[s] = L.scoef[(long long)s * L.n + i];
#pragma unroll
    for (int s = 0; s < kOctSlots; ++s) xv[s] = X(col[s]);
    c = L.cnt[i];
  }
#pragma unroll
  for (int s = 0; s < kOctSlots; ++s)
    if (s < c) acc += cf[s] * (xi - xv[s]);
  if (c > kOctSlots)
    for (int e = L.xrow[i]; e < L.xrow[i + 1]; ++e) acc += L.xcoef[e] * (xi - X(L.xcol[e]));
  return acc;
}

// Face table (count u8[6n], neighbours i32[24n]) and CSR operator of a level
// whose keys, levels and masks are already on the device and sorted.  The face
// table is returned through cnt/nbr when they are non-null, else freed.
void buildOctOperator(nbl_ctx* ctx, int maxL, const double org[3], const double fine[3], OctLevel& L,
                      DevBuf* cnt, DevBuf* nbr, hipStream_t s);
// Slot-major copy (cnt, scol, scoef, xrow, xcol, xcoef) of a level's CSR.
void buildOctSlots(nbl_ctx* ctx, OctLevel& L, hipStream_t s);
// Coefficient classes of a level with slots built (ncls = 0 when its
// coefficients take more than kOctClasses distinct values).
void buildOctClasses(nbl_ctx* ctx, OctLevel& L, hipStream_t s);

class OctMg;

class OctreeDev {
 public:
  OctreeDev(nbl_ctx* ctx, const uint64_t* keys, const uint8_t* levels, const uint8_t* masks, size_t n,
            const double origin[3], const double extent[3], int maxLevel);
  ~OctreeDev();
  size_t size() const { return L0_.n; }
  size_t nnz() const { return L0_.nnz; }
  int coefClasses() const { return L0_.ncls; }
  void permutation(uint64_t* perm);
  void faceTable(uint8_t* count, int32_t* nbr);
  // caller-order host arrays
  void apply(const double* x, double* out);
  void diagonal(double* out);
  nbl_poisson_report solve(const double* b, double* x, double tol, int maxIters);
  // Multigrid-preconditioned flexible CG on the same system (hierarchy built on
  // the first call); caller-order host arrays.  warm: x holds the initial guess.
  nbl_poisson_report solveMg(const double* b, double* x, double tol, int maxIters, bool warm = false);
  // sorted-order device arrays
  void applySorted(const double* x, double* out, hipStream_t s);
  nbl_poisson_report solveSorted(const double* b, double* x, double tol, int maxIters, hipStream_t s);
  // x0: optional initial guess (sorted; may be x itself); null starts from zero.
  nbl_poisson_report solveMgSorted(const double* b, double* x, double tol, int maxIters, hipStream_t s,
                                   const double* x0 = nullptr);
  hipStream_t stream() const { return stream_; }
  // Device time of construction (sort, validation, face table, CSR) and of the
  // last solve (ms), and a timing loop over one kernel (0: the SpMV).
  double setupMs() const { return setupMs_; }
  double lastSolveMs() const { return lastMs_; }
  void profile(int kernel, int reps, double* ms, double* bytes);
  // octreeTransportStep on phi (caller-order host arrays), `steps` times with
  // the fields resident; returns the integral sum phi_i vol_i after the last.
  double transport(const double* u, const double* v, const double* w, double* phi, double dt, int steps);
  // PhysicalSensorEvaluator::evaluate on the leaves (caller-order host arrays):
  // out[8n] field-major (vorticity, strain, Q, |grad p|, |grad u|, y+, wall
  // shear, truncation estimate), flags[3n] (near wall, reverse flow, residual
  // stagnation); speedPrev / residPrev are read and updated.  Any output may
  // be null.
  void sensors(const double* u, const double* v, const double* w, const double* p, double nu, double* speedPrev,
               double* residPrev, double* out, uint8_t* flags);
  // Multigrid hierarchy: level count, leaves per level (up to `cap` entries)
  // and the device time of its construction (ms).  Builds it if needed.
  int mgLevels(long long* counts, int cap, double* buildMs);
  // Arithmetic of the multigrid V-cycle (NBL_PREC_*): FP64 levels, FP32
  // levels, or FP32 levels with BF16 storage of level 0's pre-smoothed
  // iterate.  The flexible CG around it stays FP64.
  void setMgPrecision(int prec);
  int mgPrecision() const { return mgPrec_; }

  const OctLevel& level0() const { return L0_; }
  int maxLevel() const { return maxL_; }
  const double* origin() const { return org_; }
  const double* fine() const { return fine_; }
  Geo geo() const {
    Geo G;
    for (int a = 0; a < 3; ++a) {
      G.org[a] = org_[a];
      G.fine[a] = fine_[a];
    }
    G.maxL = maxL_;
    return G;
  }
  FaceWalk faceWalk() const {
    return FaceWalk{cnt_.as<uint8_t>(), tfirst_.as<int32_t>(), txrow_.as<int32_t>(), txnb_.as<int32_t>()};
  }
  // Face entries of the walk: the first neighbour of each direction at d*n + i
  // (also the entry of a domain-boundary face, which has none), the further
  // ones at 6n + their position in xnb.
  size_t faceEntries() const { return 6 * L0_.n + tnx_; }
  // Caller-order host array to sorted-order device array and back (on the
  // octree's stream; fromSorted synchronises it).
  void toSorted(const double* hostIn, double* dSorted);
  void fromSorted(const double* dSorted, double* hostOut);

 private:
  OctMg& mg();

  nbl_ctx* ctx_;
  int maxL_ = 0;
  double org_[3], fine_[3];
  hipStream_t stream_ = nullptr;
  OctLevel L0_;
  DevBuf perm_, cnt_, nbr_;
  DevBuf tfirst_, txrow_, txnb_;  // transport's face neighbours (first per direction slot-major, overflow)
  DevBuf sensOut_, sensFlags_;    // sensor results in caller order (allocated on first use)
  DevBuf cen_;                     // leaf centres per axis (sensors; computed on first use)
  const double* sensCen(hipStream_t s);
  size_t tnx_ = 0;
  DevBuf x_, rhs_, r_, z_, p_, Ap_, t0_, t1_;
  DevBuf ctl_, red_;
  SolveCtl* hctl_ = nullptr;
  int predicted_ = 16, predictedMg_ = 8;
  int mgPrec_ = 0;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  double setupMs_ = 0.0, lastMs_ = 0.0;
  std::unique_ptr<OctMg> mg_;
};

}  // namespace nbl

struct nbl_octree {
  nbl_ctx* ctx;
  std::unique_ptr<nbl::OctreeDev> impl;
};
