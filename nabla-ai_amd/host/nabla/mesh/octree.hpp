// Octree leaf set with the reference's construction and query surface
// (core/include/nabla/mesh/octree.hpp:27-128): a box domain, leaves addressed
// by index, uniform and box refinement with 2:1 face balance, per-leaf levels,
// masks and centres.  Leaves are kept in Morton order (the reference keeps its
// own insertion order); every index-parallel array of one tree uses the same
// order, which is all the solvers on it need.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "nbl_api.h"

namespace nabla::mesh {

enum class CellMask : uint8_t {
  Fluid = 0,
  Cut = 1,
  InsideSolid = 2,
};

struct Vec3 {
  double x{0.0};
  double y{0.0};
  double z{0.0};
};

class Octree {
 public:
  static constexpr int kMaxSupportedLevel = 21;

  // One root leaf (level 0) covering [origin, origin + extent).
  Octree(Vec3 origin, Vec3 extent, int maxLevel);
  ~Octree();
  Octree(const Octree&) = delete;
  Octree& operator=(const Octree&) = delete;

  [[nodiscard]] std::size_t cellCount() const noexcept { return level_.size(); }
  [[nodiscard]] int level(std::size_t i) const { return level_[i]; }
  [[nodiscard]] uint64_t key(std::size_t i) const { return key_[i]; }
  [[nodiscard]] CellMask mask(std::size_t i) const { return static_cast<CellMask>(mask_[i]); }
  void setMask(std::size_t i, CellMask m) { mask_[i] = static_cast<uint8_t>(m); }
  [[nodiscard]] Vec3 cellCenter(std::size_t i) const { return {centre_[3 * i], centre_[3 * i + 1], centre_[3 * i + 2]}; }
  [[nodiscard]] Vec3 origin() const { return origin_; }
  [[nodiscard]] Vec3 extent() const { return extent_; }
  [[nodiscard]] int maxLevel() const { return maxLevel_; }

  // Octree::refineUniform / refineBox (octree.cpp:360-388): every leaf, or
  // every leaf whose centre lies in [lo, hi], split once per pass, coarser
  // face neighbours first so the tree stays 2:1 balanced.  refineBox leaves
  // leaves already at maxLevel as they are.  Refinement resets the masks to
  // Fluid.
  void refineUniform(int times);
  void refineBox(Vec3 lo, Vec3 hi, int times);
  [[nodiscard]] bool isBalanced() const;

  // Leaf arrays in index order, as the solvers read them.
  [[nodiscard]] const std::vector<uint64_t>& keys() const { return key_; }
  [[nodiscard]] const std::vector<uint8_t>& levels() const { return level_; }
  [[nodiscard]] const std::vector<uint8_t>& masks() const { return mask_; }

 private:
  void refresh();

  nbl_tree* tree_ = nullptr;
  Vec3 origin_, extent_;
  int maxLevel_;
  std::vector<uint64_t> key_;
  std::vector<uint8_t> level_, mask_;
  std::vector<double> centre_;
};

}  // namespace nabla::mesh
