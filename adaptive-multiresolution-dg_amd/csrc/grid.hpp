// Host-side active-element table of the adaptive sparse grid.
//
// Elements are kept as a structure of arrays sorted by the reference hash key
// (Hash::hash_key, source/Hash.cpp:55-114).  Each element is identified by its
// per-dimension 1D slots (level-ordered cell index, see setup.hpp).  For every
// dimension d the table also carries the fiber structure the device kernels
// use instead of the reference's per-element neighbour pointer sets
// (Element::ptr_vol_alpt / ptr_flx_alpt, include/Element.h:152-155):
//   fbase[d][e]  offset of e's d-fiber in the device fiber map (fiber id << nmax)
//   fmaxl[d][e]  deepest level present in e's d-fiber
// The dense fiber map itself (fmap[d][fbase + s] = element holding slot s of
// the fiber, or -1) is scattered on the device from fbase and the slots.
// The refine / coarsen rules are the reference's (DGAdapt::refine,
// DGAdapt::coarsen, source/DGAdapt.cpp:371-404,675-688,982-1016,1250-1385)
// expressed as set operations on the table.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "setup.hpp"

namespace adg {

constexpr int kMaxDim = 4;

// Open-addressing map from packed slot vectors to element indices (linear
// probing, load factor <= 1/2): one allocation instead of one per element.
class FlatIndex {
 public:
  static constexpr uint64_t kEmpty = ~uint64_t{0};
  void reset(size_t n) {
    size_t cap = 16;
    while (cap < 2 * n + 2) cap <<= 1;
    key_.assign(cap, kEmpty);
    val_.assign(cap, -1);
    mask_ = cap - 1;
    size_ = 0;
  }
  // inserts (k, v) unless k is present; returns true when inserted
  bool insert(uint64_t k, int32_t v) {
    if (2 * (size_ + 1) > key_.size()) grow();
    for (uint64_t i = hash(k) & mask_;; i = (i + 1) & mask_) {
      if (key_[i] == kEmpty) { key_[i] = k; val_[i] = v; ++size_; return true; }
      if (key_[i] == k) return false;
    }
  }
  int32_t find(uint64_t k) const {
    if (key_.empty()) return -1;
    for (uint64_t i = hash(k) & mask_;; i = (i + 1) & mask_) {
      if (key_[i] == k) return val_[i];
      if (key_[i] == kEmpty) return -1;
    }
  }
  size_t size() const { return size_; }

 private:
  static uint64_t hash(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  void grow() {
    std::vector<uint64_t> k = std::move(key_);
    std::vector<int32_t> v = std::move(val_);
    reset(k.empty() ? 8 : k.size());
    for (size_t i = 0; i < k.size(); ++i)
      if (k[i] != kEmpty) insert(k[i], v[i]);
  }
  std::vector<uint64_t> key_;
  std::vector<int32_t> val_;
  uint64_t mask_ = 0;
  size_t size_ = 0;
};

struct HostGrid {
  int D = 0, nmax = 0;
  int64_t E = 0;
  std::vector<int32_t> slot;                // [E * D]
  std::vector<int32_t> hash;                // [E]
  FlatIndex index;
  std::vector<int32_t> dslot[kMaxDim];      // [E]  slot of element e in dim d
  std::vector<int32_t> fbase[kMaxDim];      // [E]
  std::vector<int32_t> fmaxl[kMaxDim];      // [E]
  int32_t nfib[kMaxDim] = {0, 0, 0, 0};
  int32_t maxfib[kMaxDim] = {0, 0, 0, 0};  // elements of the largest fiber

  static uint64_t pack(const int32_t* s, int D) {
    uint64_t k = 0;
    for (int d = 0; d < D; ++d) k |= static_cast<uint64_t>(s[d]) << (16 * d);
    return k;
  }
  int level(int64_t e, int d) const { return level_of_slot(slot[e * D + d]); }
  int suppt(int64_t e, int d) const { return suppt_of_slot(slot[e * D + d]); }
  int32_t find(const int32_t* s) const { return index.find(pack(s, D)); }
  int num_total_children(int64_t e) const;

  // Build the table from a list of element slot vectors (any order, no duplicates).
  void build(int D, int nmax, const std::vector<int32_t>& slots, const HashKey& hk);
  // Sparse (level sum <= level_init) or full (every level <= level_init) grid
  // as the DGSolution constructor enumerates it (source/DGSolution.cpp:10-57).
  static std::vector<int32_t> initial_slots(int D, int level_init, bool sparse);
};

// Leaves with indicator above eps get their missing children, then missing
// parents are added until no hole remains.  `norm` is the per-element
// indicator (pow(sum of squares, 0.5)).  Returns the new slot list; `added`
// counts the new elements.
std::vector<int32_t> refine_slots(const HostGrid& g, const std::vector<double>& norm, double eps, int64_t& added);

// DGAdapt::indicator_norm takes pow(sum of squares, 0.5)
// (source/DGAdapt.cpp:1018-1030).  glibc's pow and the correctly rounded sqrt
// differ by at most one ulp, so sqrt decides every comparison with the
// threshold except within a few ulps of it, where pow is evaluated: the
// refine / coarsen decisions equal those of pow everywhere.
inline double threshold_norm(double sumsq, double thr) {
  const double r = std::sqrt(sumsq);
  const double a = std::fabs(thr);
  const double band = 4.0 * (std::nextafter(a, HUGE_VAL) - a);
  return std::fabs(r - thr) <= band ? std::pow(sumsq, 1. / 2.) : r;
}

// Children-free elements whose indicator is below eta and whose maximum level
// is positive are removed repeatedly until none qualifies.
std::vector<int32_t> coarsen_slots(const HostGrid& g, const std::vector<double>& norm, double eta, int64_t& removed);

// old_of_new[i] = index in `from` of element i of `to`, or -1 when new.
std::vector<int32_t> map_new_to_old(const HostGrid& from, const HostGrid& to);

}  // namespace adg
