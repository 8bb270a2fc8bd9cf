#include "grid.hpp"

#include <algorithm>
#include <numeric>
#include <stdexcept>

namespace adg {

int HostGrid::num_total_children(int64_t e) const {
  int c = 0;
  for (int d = 0; d < D; ++d) {
    const int l = level(e, d);
    if (l == 0) c += 1;
    else if (l < nmax) c += 2;
  }
  return c;
}

void HostGrid::build(int D_, int nmax_, const std::vector<int32_t>& slots, const HashKey& hk) {
  if (D_ < 1 || D_ > kMaxDim) throw std::runtime_error("grid: dimension must be 1..4");
  if (nmax_ < 1 || nmax_ > 15) throw std::runtime_error("grid: maximum level must be 1..15");
  D = D_;
  nmax = nmax_;
  const int64_t n = static_cast<int64_t>(slots.size()) / D;
  std::vector<int32_t> keys(n);
  std::vector<int> l(D), j(D);
  for (int64_t e = 0; e < n; ++e) {
    for (int d = 0; d < D; ++d) {
      l[d] = level_of_slot(slots[e * D + d]);
      j[d] = suppt_of_slot(slots[e * D + d]);
    }
    keys[e] = hk.key(l.data(), j.data(), D);
  }
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
  E = n;
  slot.resize(n * D);
  hash.resize(n);
  index.reset(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    const int64_t e = order[i];
    for (int d = 0; d < D; ++d) slot[i * D + d] = slots[e * D + d];
    hash[i] = keys[e];
    if (i > 0 && hash[i] == hash[i - 1]) throw std::runtime_error("grid: duplicate hash key");
    index.insert(pack(&slot[i * D], D), static_cast<int32_t>(i));
  }
  // fibers per dimension: elements sorted by their slots outside dd, runs of
  // equal keys are fibers
  std::vector<uint64_t> fk(n);
  std::vector<int64_t> by(n);
  for (int dd = 0; dd < D; ++dd) {
    dslot[dd].resize(n);
    fbase[dd].resize(n);
    fmaxl[dd].resize(n);
    for (int64_t e = 0; e < n; ++e) {
      int32_t s[kMaxDim];
      for (int d = 0; d < D; ++d) s[d] = (d == dd) ? 0 : slot[e * D + d];
      fk[e] = pack(s, D);
      dslot[dd][e] = slot[e * D + dd];
    }
    std::iota(by.begin(), by.end(), 0);
    std::sort(by.begin(), by.end(), [&](int64_t a, int64_t b) { return fk[a] < fk[b]; });
    int32_t nf = 0;
    maxfib[dd] = 0;
    for (int64_t i = 0; i < n;) {
      int64_t j = i;
      int maxl = 0;
      while (j < n && fk[by[j]] == fk[by[i]]) { maxl = std::max(maxl, level_of_slot(dslot[dd][by[j]])); ++j; }
      for (int64_t q = i; q < j; ++q) {
        fbase[dd][by[q]] = nf << nmax;
        fmaxl[dd][by[q]] = maxl;
      }
      ++nf;
      maxfib[dd] = std::max<int32_t>(maxfib[dd], static_cast<int32_t>(j - i));
      i = j;
    }
    nfib[dd] = nf;
  }
}

std::vector<int32_t> HostGrid::initial_slots(int D, int level_init, bool sparse) {
  std::vector<int32_t> out;
  std::vector<int> lv(D, 0);
  while (true) {
    int sum = 0;
    for (int v : lv) sum += v;
    if (!sparse || sum <= level_init) {
      std::vector<int> jmax(D), s(D, 0);
      for (int d = 0; d < D; ++d) jmax[d] = lv[d] == 0 ? 1 : (1 << (lv[d] - 1));
      while (true) {
        for (int d = 0; d < D; ++d) out.push_back(slot_of(lv[d], 2 * s[d] + 1));
        int i = D - 1;
        while (i >= 0 && ++s[i] == jmax[i]) { s[i] = 0; --i; }
        if (i < 0) break;
      }
    }
    int i = D - 1;
    while (i >= 0 && ++lv[i] == level_init + 1) { lv[i] = 0; --i; }
    if (i < 0) break;
  }
  return out;
}

std::vector<int32_t> refine_slots(const HostGrid& g, const std::vector<double>& norm, double eps, int64_t& added) {
  const int D = g.D;
  std::vector<int32_t> out = g.slot;
  FlatIndex present;
  present.reset(static_cast<size_t>(g.E) * 2);
  for (int64_t e = 0; e < g.E; ++e) present.insert(HostGrid::pack(&g.slot[e * D], D), 0);
  std::vector<int32_t> work;  // newly added slot vectors, for the hole closure
  auto add = [&](const int32_t* s) {
    if (present.insert(HostGrid::pack(s, D), 0)) {
      out.insert(out.end(), s, s + D);
      work.insert(work.end(), s, s + D);
    }
  };
  int32_t s[kMaxDim];
  for (int64_t e = 0; e < g.E; ++e) {
    if (!(norm[e] > eps)) continue;
    // leaf: fewer existing children than possible children
    const int total = g.num_total_children(e);
    int have = 0;
    for (int d = 0; d < D; ++d) {
      for (int q = 0; q < D; ++q) s[q] = g.slot[e * D + q];
      const int sd = s[d], l = level_of_slot(sd);
      if (l == 0) { s[d] = 1; have += g.find(s) >= 0; }
      else if (l < g.nmax) {
        s[d] = 2 * sd; have += g.find(s) >= 0;
        s[d] = 2 * sd + 1; have += g.find(s) >= 0;
      }
    }
    if (have >= total) continue;
    for (int d = 0; d < D; ++d) {
      for (int q = 0; q < D; ++q) s[q] = g.slot[e * D + q];
      const int sd = s[d], l = level_of_slot(sd);
      if (l == 0) { s[d] = 1; add(s); }
      else if (l < g.nmax) { s[d] = 2 * sd; add(s); s[d] = 2 * sd + 1; add(s); }
    }
  }
  // close holes: every parent of every element must be present
  for (size_t w = 0; w < work.size(); w += D) {
    for (int d = 0; d < D; ++d) {
      for (int q = 0; q < D; ++q) s[q] = work[w + q];
      if (s[d] == 0) continue;
      s[d] = s[d] >> 1;
      add(s);
    }
  }
  added = static_cast<int64_t>(out.size() / D) - g.E;
  return out;
}

std::vector<int32_t> coarsen_slots(const HostGrid& g, const std::vector<double>& norm, double eta, int64_t& removed) {
  const int D = g.D;
  std::vector<int32_t> nchild(g.E, 0);
  std::vector<char> alive(g.E, 1);
  int32_t s[kMaxDim];
  for (int64_t e = 0; e < g.E; ++e) {
    for (int d = 0; d < D; ++d) {
      for (int q = 0; q < D; ++q) s[q] = g.slot[e * D + q];
      if (s[d] == 0) continue;
      s[d] >>= 1;
      const int32_t p = g.find(s);
      if (p >= 0) nchild[p]++;
    }
  }
  std::vector<int64_t> z;
  while (true) {
    z.clear();
    for (int64_t e = 0; e < g.E; ++e) {
      if (!alive[e] || nchild[e] != 0 || !(norm[e] < eta)) continue;
      int maxl = 0;
      for (int d = 0; d < D; ++d) maxl = std::max(maxl, g.level(e, d));
      if (maxl > 0) z.push_back(e);
    }
    if (z.empty()) break;
    for (int64_t e : z) {
      alive[e] = 0;
      for (int d = 0; d < D; ++d) {
        for (int q = 0; q < D; ++q) s[q] = g.slot[e * D + q];
        if (s[d] == 0) continue;
        s[d] >>= 1;
        const int32_t p = g.find(s);
        if (p >= 0) nchild[p]--;
      }
    }
  }
  std::vector<int32_t> out;
  out.reserve(g.slot.size());
  removed = 0;
  for (int64_t e = 0; e < g.E; ++e) {
    if (alive[e]) out.insert(out.end(), &g.slot[e * D], &g.slot[e * D] + D);
    else ++removed;
  }
  return out;
}

std::vector<int32_t> map_new_to_old(const HostGrid& from, const HostGrid& to) {
  std::vector<int32_t> m(to.E);
  for (int64_t i = 0; i < to.E; ++i) m[i] = from.find(&to.slot[i * to.D]);
  return m;
}

}  // namespace adg
