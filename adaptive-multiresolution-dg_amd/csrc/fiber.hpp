// Fiber relations of the 1D hierarchical cells, by index arithmetic.
//
// A "fiber" in dimension d is the set of elements that share (level, suppt)
// in every dimension except d.  Inside a fiber an element is identified by its
// 1D slot s (level 0 -> 0, level n>=1 -> 2^(n-1) + (j-1)/2).  The reference
// keeps, per element and dimension, explicit pointer sets of related elements
// built from interval predicates:
//   "vol": supports intersect with positive length  (Element::is_vol_alpt,
//          source/Element.cpp:265-282,380-386)  -> ancestors, self, descendants
//   "flx": supports intersect or touch, including across the periodic
//          boundary (Element::is_flx_alpt, source/Element.cpp:284-299,337-348)
// and filters them by level in FastMultiplyLU::transform_1D
// (source/FastMultiplyLU.cpp:490-491):
//   L    : source level >  target level
//   U    : source level <= target level
//   full : no filter
//
// for_each_related() enumerates exactly those source slots for a target slot,
// in ascending slot order (level-major, then position), which is the
// summation order the device kernels and the oracle both use.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define ADG_HD __host__ __device__ __forceinline__
#else
#define ADG_HD inline
#endif

namespace adg {

enum LUMode : int { LU_FULL = 0, LU_L = 1, LU_U = 2 };

ADG_HD int slot_level(int s) {
  return s == 0 ? 0 : 32 - __builtin_clz(static_cast<unsigned>(s));
}

// Calls f(source_slot) for every related source slot at levels
// [0, max_level], ascending.  flx selects the flux ("touching") relation.
template <class F>
ADG_HD void for_each_related_slots(int st, int max_level, bool flx, int lu, F&& f) {
  const int lt = slot_level(st);
  const int ls_begin = (lu == LU_L) ? lt + 1 : 0;
  const int ls_end = (lu == LU_U) ? (lt < max_level ? lt : max_level) : max_level;
  for (int ls = ls_begin; ls <= ls_end; ++ls) {
    if (ls <= 1) {  // support [0,1]: nested with every cell
      f(ls);
      continue;
    }
    const int n = 1 << (ls - 1);  // cells at level ls occupy slots [n, 2n)
    if (lt <= 1) {               // target support [0,1]: all finer cells are descendants
      for (int i = 0; i < n; ++i) f(n + i);
      continue;
    }
    const int pos = st - (1 << (lt - 1));
    if (ls <= lt) {
      const int shift = lt - ls;
      const int mask = (1 << shift) - 1;
      const int a = pos >> shift;  // the ancestor at level ls
      int lp = -1, rp = -1;
      if (flx) {
        if ((pos & mask) == 0) lp = (a - 1 + n) % n;
        if (((pos + 1) & mask) == 0) rp = ((pos + 1) >> shift) % n;
        if (lp == rp) rp = -1;
      }
      // emit {lp, a, rp} ascending
      int v[3];
      int cnt = 0;
      v[cnt++] = a;
      if (lp >= 0) v[cnt++] = lp;
      if (rp >= 0) v[cnt++] = rp;
      for (int i = 1; i < cnt; ++i)
        for (int j = i; j > 0 && v[j - 1] > v[j]; --j) { int t = v[j]; v[j] = v[j - 1]; v[j - 1] = t; }
      for (int i = 0; i < cnt; ++i) f(n + v[i]);
    } else {
      const int shift = ls - lt;
      const int lo = pos << shift, hi = (pos + 1) << shift;
      if (flx) {
        const int lp = (lo - 1 + n) % n, rp = hi % n;
        if (hi == n) f(n + rp);   // wrapped right neighbour is cell 0
        if (lo > 0) f(n + lp);
        for (int i = lo; i < hi; ++i) f(n + i);
        if (hi < n) f(n + rp);
        if (lo == 0) f(n + lp);   // wrapped left neighbour is the last cell
      } else {
        for (int i = lo; i < hi; ++i) f(n + i);
      }
    }
  }
}

}  // namespace adg

namespace adg {

// The same enumeration as maximal runs of consecutive slots: f(lo, hi) is
// called for [lo, hi) ranges in ascending order.  Device kernels use the runs
// to issue the loads of several sources before accumulating them in order.
template <class F>
ADG_HD void for_each_related_ranges(int st, int max_level, bool flx, int lu, F&& f) {
  int plo = -1, phi = -1;
  for_each_related_slots(st, max_level, flx, lu, [&](int s) {
    if (s == phi) { ++phi; return; }
    if (plo >= 0) f(plo, phi);
    plo = s;
    phi = s + 1;
  });
  if (plo >= 0) f(plo, phi);
}

// Slot-by-slot form built on the runs (used by the relation tests).
template <class F>
ADG_HD void for_each_related(int st, int max_level, bool flx, int lu, F&& f) {
  for_each_related_ranges(st, max_level, flx, lu, [&](int lo, int hi) {
    for (int s = lo; s < hi; ++s) f(s);
  });
}

// Present source slots of a target, ascending.  Short runs (ancestors and flux
// neighbours) are passed on as they are and resolved through the fiber map;
// longer runs (descendants) are read from the fiber's occupancy bits, 32
// slots per load, so that absent cells cost nothing.
template <class F>
ADG_HD void for_each_present(int st, int maxl, bool flx, int lu, const uint32_t* bits, int fb, F&& f) {
  for_each_related_ranges(st, maxl, flx, lu, [&](int lo, int hi) {
    if (hi - lo <= 2) {
      for (int s = lo; s < hi; ++s) f(s);
      return;
    }
    const int b0 = fb + lo, b1 = fb + hi;  // bit range [b0, b1)
    for (int w = b0 >> 5; w <= (b1 - 1) >> 5; ++w) {
      uint32_t m = bits[w];
      const int wb = w << 5;
      if (wb < b0) m &= ~0u << (b0 - wb);
      if (wb + 32 > b1) m &= ~0u >> (wb + 32 - b1);
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        f(wb + b - fb);
      }
    }
  });
}

}  // namespace adg
