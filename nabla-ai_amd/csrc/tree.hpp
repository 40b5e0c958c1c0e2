// tree.hpp — host-side builder of a 2:1-balanced linear octree (set-up).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace nbl {

class HostTree {
 public:
  HostTree(const double origin[3], const double extent[3], int maxLevel);
  size_t size() const { return key_.size(); }
  int maxLevel() const { return maxL_; }
  void refineUniform(int times);
  // Balanced refinement of the flagged leaves (index = current sorted order).
  void refineFlagged(const uint8_t* flags, size_t n);
  bool balanced() const;
  void leaves(uint64_t* keys, uint8_t* levels, double* centers) const;

 private:
  long long find(uint64_t k) const;
  long long coarseNb(size_t i, int dir, int& nl) const;
  void splitMarked(const std::vector<uint8_t>& mark);

  int maxL_;
  double org_[3], fine_[3];
  std::vector<uint64_t> key_;
  std::vector<uint8_t> lev_;
};

}  // namespace nbl

struct nbl_tree {
  explicit nbl_tree(const double o[3], const double e[3], int ml) : impl(o, e, ml) {}
  nbl::HostTree impl;
};
