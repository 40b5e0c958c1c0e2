// tree.hip — host-side construction of a 2:1-balanced linear octree (the
// set-up step of the octree path; reference core/src/mesh/octree.cpp:288-388,
// 544-560).
//
// Design: leaves live in one array sorted by Morton key (children of a split
// leaf are emitted in Morton order at their parent's position, so the array
// stays sorted without re-sorting).  Neighbour probes are binary searches.
// A refinement request splits the flagged leaves together with the balance
// closure computed by a worklist: a leaf at level L may split only when no
// face neighbour is coarser than L, so any such neighbour is marked too, and
// passes repeat until every requested split is done and the tree is balanced.
// The result is the minimal balanced refinement containing the requested
// splits, i.e. the same leaf set the reference's recursive refineByMorton
// produces.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "tree.hpp"

namespace nbl {
namespace {

uint64_t spread3h(uint32_t a) {
  uint64_t x = a & 0x1fffffull;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
uint32_t gather3h(uint64_t x) {
  x &= 0x1249249249249249ull;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
  x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
  x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
  x = (x ^ (x >> 32)) & 0x1fffffull;
  return uint32_t(x);
}
uint64_t mencH(uint32_t x, uint32_t y, uint32_t z) { return spread3h(x) | (spread3h(y) << 1) | (spread3h(z) << 2); }

}  // namespace

HostTree::HostTree(const double origin[3], const double extent[3], int maxLevel) : maxL_(maxLevel) {
  if (maxLevel < 1 || maxLevel > 21) throw std::invalid_argument("Octree: maxLevel must be in [1, 21]");
  if (!(extent[0] > 0 && extent[1] > 0 && extent[2] > 0)) throw std::invalid_argument("Octree: extent must be positive");
  const double inv = 1.0 / double(1u << maxLevel);
  for (int a = 0; a < 3; ++a) {
    org_[a] = origin[a];
    fine_[a] = extent[a] * inv;
  }
  key_.assign(1, 0);
  lev_.assign(1, 0);
}

long long HostTree::find(uint64_t k) const {
  auto it = std::lower_bound(key_.begin(), key_.end(), k);
  return (it != key_.end() && *it == k) ? (long long)(it - key_.begin()) : -1;
}

// Same-or-coarser leaf across face `dir` of leaf i, -1 at the domain boundary
// or when the neighbour is finer (octree.cpp:201-227).
long long HostTree::coarseNb(size_t i, int dir, int& nl) const {
  const uint64_t k = key_[i];
  const int L = lev_[i];
  const int axis = dir >> 1;
  const uint32_t a[3] = {gather3h(k), gather3h(k >> 1), gather3h(k >> 2)};
  int64_t q[3] = {a[0], a[1], a[2]};
  q[axis] += (dir & 1) ? int64_t(1u << (maxL_ - L)) : -int64_t(1u << (maxL_ - L));
  if (q[axis] < 0 || q[axis] >= (int64_t(1) << maxL_)) return -1;
  for (int l = L; l >= 0; --l) {
    const uint32_t m = ~((1u << (maxL_ - l)) - 1u);
    const long long hit = find(mencH(uint32_t(q[0]) & m, uint32_t(q[1]) & m, uint32_t(q[2]) & m));
    if (hit >= 0 && lev_[hit] == l) {
      nl = l;
      return hit;
    }
  }
  return -1;
}

void HostTree::splitMarked(const std::vector<uint8_t>& mark) {
  std::vector<uint64_t> k2;
  std::vector<uint8_t> l2;
  size_t extra = 0;
  for (uint8_t m : mark) extra += m ? 7 : 0;
  k2.reserve(key_.size() + extra);
  l2.reserve(key_.size() + extra);
  for (size_t i = 0; i < key_.size(); ++i) {
    if (!mark[i]) {
      k2.push_back(key_[i]);
      l2.push_back(lev_[i]);
      continue;
    }
    const int L = lev_[i] + 1;
    const uint64_t step = spread3h(1u << (maxL_ - L));  // x offset of one child
    // children c = 0..7 (x = c&1, y = c>>1&1, z = c>>2&1, octree.cpp:307-311)
    // have keys parent + c * step in Morton order
    for (uint64_t c = 0; c < 8; ++c) {
      k2.push_back(key_[i] + c * step);
      l2.push_back(uint8_t(L));
    }
  }
  key_.swap(k2);
  lev_.swap(l2);
}

void HostTree::refineUniform(int times) {
  for (int t = 0; t < times; ++t) {
    for (uint8_t l : lev_)
      if (l >= maxL_) throw std::runtime_error("refine: cell already at maxLevel");
    splitMarked(std::vector<uint8_t>(key_.size(), 1));
  }
}

void HostTree::refineFlagged(const uint8_t* flags, size_t n) {
  if (n != key_.size()) throw std::invalid_argument("refine: flag count differs from leaf count");
  // pending splits as (key, level): a pending leaf splits in a pass once none
  // of its face neighbours is coarser; otherwise those neighbours become
  // pending too and it waits for the next pass
  std::vector<uint64_t> wk;
  std::vector<uint8_t> wl;
  for (size_t i = 0; i < n; ++i)
    if (flags[i]) {
      if (lev_[i] >= maxL_) throw std::runtime_error("refine: cell already at maxLevel");
      wk.push_back(key_[i]);
      wl.push_back(lev_[i]);
    }
  while (!wk.empty()) {
    std::vector<uint8_t> mark(key_.size(), 0), blocked(key_.size(), 0);
    std::vector<size_t> work;
    for (size_t q = 0; q < wk.size(); ++q) {
      const long long i = find(wk[q]);
      if (i >= 0 && lev_[i] == wl[q] && !mark[i]) {
        mark[i] = 1;
        work.push_back(size_t(i));
      }
    }
    while (!work.empty()) {
      const size_t i = work.back();
      work.pop_back();
      for (int d = 0; d < 6; ++d) {
        int nl = 0;
        const long long j = coarseNb(i, d, nl);
        if (j >= 0 && nl < lev_[i]) {
          blocked[i] = 1;
          if (!mark[j]) {
            mark[j] = 1;
            work.push_back(size_t(j));
          }
        }
      }
    }
    std::vector<uint8_t> now(key_.size(), 0);
    wk.clear();
    wl.clear();
    for (size_t i = 0; i < key_.size(); ++i) {
      if (!mark[i]) continue;
      if (blocked[i]) {
        wk.push_back(key_[i]);
        wl.push_back(lev_[i]);
      } else {
        now[i] = 1;
      }
    }
    splitMarked(now);
  }
}

bool HostTree::balanced() const {
  for (size_t i = 0; i < key_.size(); ++i)
    for (int d = 0; d < 6; ++d) {
      int nl = 0;
      if (coarseNb(i, d, nl) >= 0 && lev_[i] - nl > 1) return false;
    }
  return true;
}

void HostTree::leaves(uint64_t* keys, uint8_t* levels, double* centers) const {
  for (size_t i = 0; i < key_.size(); ++i) {
    if (keys) keys[i] = key_[i];
    if (levels) levels[i] = lev_[i];
    if (centers) {
      const uint32_t sz = 1u << (maxL_ - lev_[i]);
      for (int a = 0; a < 3; ++a) {
        const double units = double(gather3h(key_[i] >> a)) + 0.5 * double(sz);
        centers[3 * i + a] = org_[a] + units * fine_[a];
      }
    }
  }
}

}  // namespace nbl
