// Refine and coarsen of the union element table on the device.
//
// The rules are the reference's (DGAdapt::refine and DGAdapt::coarsen,
// source/DGAdapt.cpp:371-404,675-688,982-1016,1250-1385) in the form the
// host code states them (grid.cpp: refine_slots, coarsen_slots,
// HostGrid::build):
//   refine   every leaf (fewer children present than possible) whose norm
//            exceeds eps gets its missing children; then every missing parent
//            of a new element is added until no hole remains;
//   coarsen  childless elements of positive maximum level whose norm is below
//            eta are removed, all at once, repeatedly until none qualifies;
//   table    elements ordered by (instance, Hash::hash_key); fibers of
//            dimension d ordered by (instance, slots outside d), fiber f of
//            the union at fbase = f << NMAX, fmaxl the deepest level in it.
// Element sets, order and fiber tables equal the host build's (checked by
// Context::check_device_grid).  New elements are identified by a 64-bit key:
// instance above D fields of NMAX bits, field d = the slot in dimension d.
// Lookups in the current table go through its fiber maps (children and
// parents lie in an element's fibers) or a binary search in its sorted
// (instance << 32 | hash key) array; only new keys enter a hash table.
//
// The norm test takes sqrt of the sum of squares; within four ulps of the
// threshold the reference's pow(s, 0.5) decides, which only the host has, so
// such a norm makes refine_device / coarsen_device return -1 before anything
// changes, and the host path runs instead (norm_flags_dev instead sends just
// those elements to the host).  A refinement larger than its new-element list
// runs again with a larger list.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

#include "context.hpp"

namespace adg {
namespace {

constexpr int kAB = 256;
inline unsigned ablocks(int64_t n) { return static_cast<unsigned>(std::max<int64_t>((n + kAB - 1) / kAB, 1)); }
constexpr uint64_t kFree = ~uint64_t{0};
constexpr int64_t kMaxNew = int64_t{1} << 28;  // largest new-element list

// counter words: [0] new (refine) or removed (coarsen) elements, [1] flags,
// [2] coarsen candidates, [3] first closure layer, [4 + t] elements removed
// by coarsen pass t
enum : int32_t { AF_BAND = 1, AF_FULL = 2, AF_DUP = 4, AF_OPEN = 8, AF_NEG = 16 };
constexpr int kCtr = 4 + 4 * 15 + 2;

__device__ inline int slot_level(int s) { return s == 0 ? 0 : 32 - __clz(s); }

__device__ inline uint64_t amix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline int field(uint64_t key, int d, int nb) {
  return static_cast<int>((key >> (nb * d)) & ((uint64_t{1} << nb) - 1));
}

__device__ inline uint64_t with_field(uint64_t key, int d, int nb, int v) {
  const uint64_t m = ((uint64_t{1} << nb) - 1) << (nb * d);
  return (key & ~m) | (static_cast<uint64_t>(v) << (nb * d));
}

// insert-if-absent with bounded probing: 1 inserted (slot in *at), 0 present,
// -1 full.  At load <= 1/2 linear probing ends within a few steps; a table
// filled past that (a refinement larger than the new-element list) gives up
// after kMaxProbe steps and the call runs again with a larger table.
constexpr uint64_t kMaxProbe = 512;
__device__ inline int tab_insert(uint64_t* hkey, uint64_t mask, uint64_t key, uint64_t* at) {
  uint64_t h = amix(key) & mask;
  for (uint64_t p = 0; p <= mask && p < kMaxProbe; ++p) {
    const unsigned long long prev =
        atomicCAS(reinterpret_cast<unsigned long long*>(hkey + h), static_cast<unsigned long long>(kFree),
                  static_cast<unsigned long long>(key));
    if (prev == kFree) { *at = h; return 1; }
    if (prev == key) return 0;
    h = (h + 1) & mask;
  }
  return -1;
}

// sort key (instance, Hash::hash_key) of an element: the support part of the
// key in mixed radix 2^level plus the level-vector part (source/Hash.cpp:55-114)
__device__ inline uint64_t sort_key(const int* s, int in, int D, const int32_t* lvtab, int32_t* ctr) {
  int ind1 = 0, code = 0;
  for (int t = 0; t < D; ++t) {
    const int l = slot_level(s[t]);
    ind1 = ind1 * (1 << l) + (l == 0 ? 0 : s[t] - (1 << (l - 1)));
    code = code * 16 + l;
  }
  const int h = ind1 + lvtab[code];
  if (h < 0) atomicOr(ctr + 1, AF_NEG);
  return (static_cast<uint64_t>(in) << 32) | static_cast<uint32_t>(h);
}

// whether the sorted keys of the current table hold k (binary search)
__device__ inline bool old_has(const uint64_t* ukey, int64_t n, uint64_t k) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ukey[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo < n && ukey[lo] == k;
}

// DGAdapt::indicator_norm's pow(s, 0.5) compared with a threshold: false
// with r = sqrt(s) outside the band where sqrt and pow may round apart
__device__ inline bool in_band(double s, double thr, double& r) {
  r = sqrt(s);
  const double a = fabs(thr);
  const double band = 4.0 * (nextafter(a, __builtin_huge_val()) - a);
  return fabs(r - thr) <= band;
}

struct AdArgs {
  const int32_t* slot[kMaxDim];
  const int32_t* fbase[kMaxDim];
  const int32_t* fmap[kMaxDim];
  const int32_t* inst;
  const double* sumsq;
  const double* thr;  // per instance
  const int32_t* flags;  // refine: elements to refine (if set; else the norm test against thr)
  const uint64_t* ukey;   // sorted (instance, hash key) of the current table
  const int32_t* lvtab;   // level-vector part of the hash key
  uint64_t* hkey;         // insert-if-absent table of the new elements
  uint64_t mask;
  uint64_t* newkey;
  int32_t* ctr;
  int32_t* nchild;
  int32_t* zmark;
  int32_t* dec;
  int32_t* cand;  // coarsen candidates (below eta, positive maximum level)
  int64_t E;
  int D, nb, nmax;
  int32_t cap_new;
};

__device__ inline uint64_t elem_key(const AdArgs& a, int64_t e) {
  uint64_t k = static_cast<uint64_t>(a.inst[e]) << (a.D * a.nb);
  for (int d = 0; d < a.D; ++d) k |= static_cast<uint64_t>(a.slot[d][e]) << (a.nb * d);
  return k;
}

__device__ inline void add_new(const AdArgs& a, uint64_t key) {
  uint64_t at;
  const int r = tab_insert(a.hkey, a.mask, key, &at);
  if (r < 0) { atomicOr(a.ctr + 1, AF_FULL); return; }
  if (r == 0) return;
  const int32_t idx = atomicAdd(a.ctr, 1);
  if (idx < a.cap_new) a.newkey[idx] = key;
  else atomicOr(a.ctr + 1, AF_FULL);
}

__global__ __launch_bounds__(kAB) void k_ad_fill64(uint64_t* p, int64_t n, uint64_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kAB) void k_ad_fill32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// an empty fiber map: every entry -1, every occupancy word 0
__global__ __launch_bounds__(kAB) void k_ad_clear_map(int32_t* fmap, int64_t nmap, uint32_t* fbits, int64_t nwords) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nmap) fmap[i] = -1;
  if (i < nwords) fbits[i] = 0u;
}

// refine: leaves above eps add their missing children.  The leaf test and
// the presence of each child read the fiber maps of the current grid
// (children of e in dimension d lie in e's d-fiber), so they see the grid
// before this refinement, as the host's do; the table dedups the new ones.
__global__ __launch_bounds__(kAB) void k_ad_children(AdArgs a) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  if (a.flags) {
    if (!a.flags[e]) return;
  } else {
    double r;
    const double t = a.thr[a.inst[e]];
    if (in_band(a.sumsq[e], t, r)) { atomicOr(a.ctr + 1, AF_BAND); return; }
    if (!(r > t)) return;
  }
  int s[kMaxDim];
  int total = 0, have = 0;
  for (int d = 0; d < a.D; ++d) {
    s[d] = a.slot[d][e];
    const int l = slot_level(s[d]);
    const int32_t* fm = a.fmap[d] + a.fbase[d][e];
    if (l == 0) {
      total += 1;
      have += fm[1] >= 0;
    } else if (l < a.nmax) {
      total += 2;
      have += (fm[2 * s[d]] >= 0) + (fm[2 * s[d] + 1] >= 0);
    }
  }
  if (have >= total) return;
  // the missing ones (present children are in the fiber map)
  const uint64_t key = elem_key(a, e);
  for (int d = 0; d < a.D; ++d) {
    const int l = slot_level(s[d]);
    const int32_t* fm = a.fmap[d] + a.fbase[d][e];
    if (l == 0) {
      if (fm[1] < 0) add_new(a, with_field(key, d, a.nb, 1));
    } else if (l < a.nmax) {
      if (fm[2 * s[d]] < 0) add_new(a, with_field(key, d, a.nb, 2 * s[d]));
      if (fm[2 * s[d] + 1] < 0) add_new(a, with_field(key, d, a.nb, 2 * s[d] + 1));
    }
  }
}

// One element of a hole-closure layer: its parents that neither the current
// table (binary search in its sorted keys) nor the new-key table holds.
__device__ inline void close_one(const AdArgs& a, int32_t w) {
  const uint64_t key = a.newkey[w];
  int s[kMaxDim];
  for (int d = 0; d < a.D; ++d) s[d] = field(key, d, a.nb);
  const int in = static_cast<int>(key >> (a.D * a.nb));
  for (int d = 0; d < a.D; ++d) {
    if (s[d] == 0) continue;
    const int sd = s[d];
    s[d] = sd >> 1;
    const bool have = old_has(a.ukey, a.E, sort_key(s, in, a.D, a.lvtab, a.ctr));
    s[d] = sd;
    if (!have) add_new(a, with_field(key, d, a.nb, sd >> 1));
  }
}

// ctr[3] = the children's count, the extent of the first closure layer
__global__ void k_ad_snap(int32_t* ctr, int32_t cap_new) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[3] = min(ctr[0], cap_new);
}

// first closure layer (the parents of the children, the largest), grid-stride
__global__ __launch_bounds__(kAB) void k_ad_close_first(AdArgs a) {
  const int32_t n = a.ctr[3];
  for (int32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x) close_one(a, w);
}

// The remaining layers in one workgroup: layer after layer, the parents of the
// elements the previous layer added.  Every layer lowers the level sum by one,
// so `layers` = D * NMAX layers after the first always suffice; a last layer
// that still added something is flagged.
constexpr int kCloseThreads = 1024;
__global__ __launch_bounds__(kCloseThreads) void k_ad_close(AdArgs a, int layers) {
  __shared__ int32_t range[2];
  if (threadIdx.x == 0) {
    range[0] = a.ctr[3];
    range[1] = min(atomicAdd(a.ctr, 0), a.cap_new);
  }
  __syncthreads();
  for (int it = 0; it < layers; ++it) {
    const int32_t b = range[0], en = range[1];
    if (b >= en) break;
    for (int32_t w = b + static_cast<int32_t>(threadIdx.x); w < en; w += kCloseThreads) close_one(a, w);
    __syncthreads();
    if (threadIdx.x == 0) {
      range[0] = en;
      range[1] = min(atomicAdd(a.ctr, 0), a.cap_new);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && range[0] != range[1]) atomicOr(a.ctr + 1, AF_OPEN);
}

// the union table after a refinement: old elements, then the new list
struct UnionArgs {
  const int32_t* slot[kMaxDim];
  const int32_t* inst;
  const uint64_t* newkey;
  const int32_t* lvtab;  // level-vector part of the hash key
  uint64_t* keys;
  int32_t* vals;
  int32_t* ctr;
  int64_t E_old, N;
  int D, nb;
};

__device__ inline void union_elem(const UnionArgs& a, int64_t q, int* s, int& in) {
  if (q < a.E_old) {
    for (int d = 0; d < a.D; ++d) s[d] = a.slot[d][q];
    in = a.inst[q];
  } else {
    const uint64_t key = a.newkey[q - a.E_old];
    for (int d = 0; d < a.D; ++d) s[d] = field(key, d, a.nb);
    in = static_cast<int>(key >> (a.D * a.nb));
  }
}

// sort key (instance, Hash::hash_key): the support part of the key in mixed
// radix 2^level plus the level-vector part (source/Hash.cpp:55-114)
__global__ __launch_bounds__(kAB) void k_ad_union_keys(UnionArgs a) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.N) return;
  int s[kMaxDim], in;
  union_elem(a, q, s, in);
  a.keys[q] = sort_key(s, in, a.D, a.lvtab, a.ctr);
  a.vals[q] = static_cast<int32_t>(q);
}

struct OutArgs {
  int32_t* slot[kMaxDim];
  int32_t* inst;
  int32_t* lsum;
  int32_t* remap;  // old index of every new element, -1 when new
  uint64_t* ukey;  // sort key of every element
  int B;
};

__device__ inline void put_elem(const OutArgs& o, int64_t q, const int* s, int in, int32_t old, uint64_t key, int D) {
  int ls = 0;
  for (int d = 0; d < D; ++d) {
    o.slot[d][q] = s[d];
    ls += slot_level(s[d]);
  }
  o.inst[q] = in;
  o.lsum[q] = ls;
  o.remap[q] = old;
  o.ukey[q] = key;
}

__global__ __launch_bounds__(kAB) void k_ad_union_gather(UnionArgs a, const uint64_t* skeys, const int32_t* svals,
                                                         OutArgs o) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.N) return;
  if (q > 0 && skeys[q - 1] == skeys[q]) atomicOr(a.ctr + 1, AF_DUP);
  const int32_t src = svals[q];
  int s[kMaxDim], in;
  union_elem(a, src, s, in);
  put_elem(o, q, s, in, src < a.E_old ? src : -1, skeys[q], a.D);
}

// coarsen: children of every element, then removal passes
__global__ __launch_bounds__(kAB) void k_ad_nchild(AdArgs a) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  // the parent in dimension d lies in e's d-fiber
  for (int d = 0; d < a.D; ++d) {
    const int s = a.slot[d][e];
    if (s == 0) continue;
    const int32_t p = a.fmap[d][a.fbase[d][e] + (s >> 1)];
    if (p >= 0) atomicAdd(a.nchild + p, 1);
  }
}

// (first_pass: also mark pass 1, the childless candidates, into list0 with
// its count in ctr[3]; the child counts are final here)
__global__ __launch_bounds__(kAB) void k_ad_below(AdArgs a, int32_t* list0) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  bool any = false;
  for (int d = 0; d < a.D; ++d) any = any || a.slot[d][e] != 0;
  double r;
  const double t = a.thr[a.inst[e]];
  const bool band = in_band(a.sumsq[e], t, r);
  if (band && any) atomicOr(a.ctr + 1, AF_BAND);
  const bool below = any && !band && r < t;
  a.dec[e] = below ? 1 : 0;
  a.zmark[e] = 0;
  // only candidates can be removed: the passes walk this list (its order
  // does not matter, every pass decides on the state before it)
  if (below) a.cand[atomicAdd(a.ctr + 2, 1)] = static_cast<int32_t>(e);
  if (below && list0 && a.nchild[e] == 0) {
    a.zmark[e] = 1;
    list0[atomicAdd(a.ctr + 3, 1)] = static_cast<int32_t>(e);
  }
}

// pass `it` (1, 2, ...): every element alive, childless and below eta is
// marked from the state after the previous pass
// (a pass after one that removed nothing changes nothing and exits)
__global__ __launch_bounds__(kAB) void k_ad_mark(AdArgs a, int it, int last) {
  if (it > 1 && a.ctr[4 + it - 1] == 0) return;
  const int32_t n = a.ctr[2];
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t e = a.cand[j];
    if (a.zmark[e] != 0 || a.nchild[e] != 0) continue;
    a.zmark[e] = it;
    atomicAdd(a.ctr, 1);
    atomicAdd(a.ctr + 4 + it, 1);
    if (last) atomicOr(a.ctr + 1, AF_OPEN);
  }
}

__global__ __launch_bounds__(kAB) void k_ad_unlink(AdArgs a, int it) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a.ctr[4 + it] == 0) return;
  const int32_t n = a.ctr[2];
  for (int64_t j = e; j < n; j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t c = a.cand[j];
    if (a.zmark[c] != it) continue;
    for (int d = 0; d < a.D; ++d) {
      const int s = a.slot[d][c];
      if (s == 0) continue;
      const int32_t p = a.fmap[d][a.fbase[d][c] + (s >> 1)];
      if (p >= 0) atomicSub(a.nchild + p, 1);
    }
  }
}

__global__ __launch_bounds__(kAB) void k_ad_drop(AdArgs a, const int32_t* drop) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < a.E) a.zmark[e] = drop[a.inst[e]] ? 1 : 0;
}

// All removal passes in one workgroup (tables of up to kSmallTable
// elements), the passes of k_ad_mark / k_ad_unlink with barriers between
// them.  Pass 1 is the childless candidates; after it an element can only become
// removable when its last child goes, so the thread whose decrement takes a
// parent's count to zero puts that parent (if it is a candidate) on the next
// pass's list.  Stops after a pass that removes nothing.
constexpr int64_t kSmallTable = int64_t{1} << 17;
__global__ __launch_bounds__(kCloseThreads) void k_ad_passes(AdArgs a, int passes, int32_t* list0, int32_t* list1) {
  __shared__ int32_t count[2];
  // pass 1 was marked by k_ad_below (list0, ctr[3] elements)
  if (threadIdx.x == 0) {
    count[0] = a.ctr[3];
    count[1] = 0;
  }
  __syncthreads();
  int32_t* cur = list0;
  int32_t* nxt = list1;
  for (int it = 1; it <= passes; ++it) {
    const int32_t m = count[(it - 1) & 1];
    if (m == 0) break;
    if (threadIdx.x == 0) {
      atomicAdd(a.ctr, m);
      if (it == passes) atomicOr(a.ctr + 1, AF_OPEN);
      count[it & 1] = 0;
    }
    if (it == passes) break;
    __syncthreads();
    for (int32_t j = threadIdx.x; j < m; j += kCloseThreads) {
      const int32_t e = cur[j];
      for (int d = 0; d < a.D; ++d) {
        const int s = a.slot[d][e];
        if (s == 0) continue;
        const int32_t p = a.fmap[d][a.fbase[d][e] + (s >> 1)];
        if (p >= 0 && atomicSub(a.nchild + p, 1) == 1 && a.dec[p] && a.zmark[p] == 0) {
          a.zmark[p] = it + 1;
          nxt[atomicAdd(&count[it & 1], 1)] = p;
        }
      }
    }
    __syncthreads();
    int32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
}

__global__ __launch_bounds__(kAB) void k_ad_keep(const int32_t* zmark, int32_t* keep, int64_t E) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < E) keep[e] = zmark[e] == 0 ? 1 : 0;
}

// survivors keep their order (the table stays sorted by instance and key)
__global__ __launch_bounds__(kAB) void k_ad_compact(AdArgs a, const int32_t* pos, OutArgs o) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E || a.zmark[e] != 0) return;
  int s[kMaxDim];
  for (int d = 0; d < a.D; ++d) s[d] = a.slot[d][e];
  put_elem(o, pos[e], s, a.inst[e], static_cast<int32_t>(e), a.ukey[e], a.D);
}

// fiber tables of dimension dd
struct FibArgs {
  const int32_t* slot[kMaxDim];
  const int32_t* inst;
  uint64_t* keys;
  int32_t* vals;
  int64_t N;
  int D, nb, dd;
};

__global__ __launch_bounds__(kAB) void k_ad_fib_keys(FibArgs a) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.N) return;
  uint64_t k = static_cast<uint64_t>(a.inst[q]) << (a.D * a.nb);
  for (int d = 0; d < a.D; ++d)
    if (d != a.dd) k |= static_cast<uint64_t>(a.slot[d][q]) << (a.nb * d);
  a.keys[q] = k;
  a.vals[q] = static_cast<int32_t>(q);
}

// run starts of the sorted fiber keys; zeroes the per-run accumulators
__global__ __launch_bounds__(kAB) void k_ad_run_flags(const uint64_t* skeys, int32_t* flags, int64_t N, int32_t* runmax,
                                                     int32_t* runlen) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  flags[i] = (i == 0 || skeys[i] != skeys[i - 1]) ? 1 : 0;
  runmax[i] = 0;
  runlen[i] = 0;
}

__global__ __launch_bounds__(kAB) void k_ad_runs(const int32_t* svals, const int32_t* run, const int32_t* slot_dd,
                                                int32_t* runmax, int32_t* runlen, int32_t* fbase, int64_t N, int nmax,
                                                int32_t* nfib) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int32_t e = svals[i];
  const int32_t r = run[i] - 1;
  fbase[e] = r << nmax;
  atomicMax(runmax + r, slot_level(slot_dd[e]));
  atomicAdd(runlen + r, 1);
  if (i == N - 1) *nfib = r + 1;
}

// fiber levels; at the first element of every fiber, the largest fiber and
// the instance's deepest level in dimension dd (one atomic per fiber)
__global__ __launch_bounds__(kAB) void k_ad_run_levels(const int32_t* svals, const int32_t* run, const int32_t* runmax,
                                                      const int32_t* runlen, int32_t* fmaxl, int64_t N,
                                                      int32_t* maxfib, const int32_t* inst, int32_t* lmax, int D,
                                                      int dd) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int32_t r = run[i] - 1;
  fmaxl[svals[i]] = runmax[r];
  if (i == 0 || run[i] != run[i - 1]) {
    atomicMax(maxfib, runlen[r]);
    atomicMax(lmax + inst[svals[i]] * D + dd, runmax[r]);
  }
}

// first and last element of every instance in the sorted table
__global__ __launch_bounds__(kAB) void k_ad_bounds(const int32_t* inst, int64_t N, int32_t* first, int32_t* last) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= N) return;
  const int32_t in = inst[q];
  if (q == 0 || inst[q - 1] != in) first[in] = static_cast<int32_t>(q);
  if (q == N - 1 || inst[q + 1] != in) last[in] = static_cast<int32_t>(q);
}

__global__ __launch_bounds__(kAB) void k_ad_fmap(int32_t* fmap, uint32_t* fbits, const int32_t* fbase,
                                                const int32_t* slot, int64_t N) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= N) return;
  const int b = fbase[e] + slot[e];
  fmap[b] = static_cast<int32_t>(e);
  atomicOr(&fbits[b >> 5], 1u << (b & 31));
}

// whether each element is a leaf (fewer children present than possible)
__global__ __launch_bounds__(kAB) void k_ad_leaf(AdArgs a, int32_t* leaf) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  int total = 0, have = 0;
  for (int d = 0; d < a.D; ++d) {
    const int s = a.slot[d][e];
    const int l = slot_level(s);
    const int32_t* fm = a.fmap[d] + a.fbase[d][e];
    if (l == 0) {
      total += 1;
      have += fm[1] >= 0;
    } else if (l < a.nmax) {
      total += 2;
      have += (fm[2 * s] >= 0) + (fm[2 * s + 1] >= 0);
    }
  }
  leaf[e] = have < total ? 1 : 0;
}

// a per-element flag carried to the table after a refinement (new: 0)
__global__ __launch_bounds__(kAB) void k_ad_carry(const int32_t* remap, const int32_t* old, int32_t* out, int64_t N) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < N) out[q] = remap[q] >= 0 ? old[remap[q]] : 0;
}

__global__ __launch_bounds__(kAB) void k_ad_prob_flags(const float* prob, double thr, int32_t* flags, int64_t E) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < E) flags[e] = static_cast<double>(prob[e]) > thr ? 1 : 0;
}

// flags[e] = (mask[e] if a mask) and norm above the instance's threshold;
// norms within the band where pow decides are listed for the host
__global__ __launch_bounds__(kAB) void k_ad_norm_flags(AdArgs a, const int32_t* mask, int32_t* flags, int32_t* band) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  if (mask && !mask[e]) {
    flags[e] = 0;
    return;
  }
  double r;
  const double t = a.thr[a.inst[e]];
  if (in_band(a.sumsq[e], t, r)) {
    flags[e] = 0;
    band[atomicAdd(a.ctr + 2, 1)] = static_cast<int32_t>(e);
    return;
  }
  flags[e] = r > t ? 1 : 0;
}

__global__ void k_ad_set_flags(const int32_t* ids, const int32_t* vals, int32_t n, int32_t* flags) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[ids[i]] = vals[i];
}

int bits_for(int64_t n) {  // bits to hold 0..n-1 (at least one)
  int b = 1;
  while ((int64_t{1} << b) < n) ++b;
  return b;
}

template <class T>
void to_device(DevBuf<T>& b, const std::vector<T>& h, hipStream_t s) {
  b.ensure(std::max<size_t>(h.size(), 1));
  if (!h.empty()) hip_check(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

}  // namespace

bool Context::device_adapt_fits(int64_t n) const {
  if (!device_adapt || recording || !up_host.empty() || n <= 0 || B() < 1) return false;
  // element keys in 63 bits, int32 element indices (the fiber bases are
  // checked against int32 once the fibers are counted)
  return D * nmax + bits_for(B()) <= 62 && n < (int64_t{1} << 31) - 1;
}

// Radix sort of n (key, value) pairs between the two buffers of each array
// (the double-buffer form: no copy back); afterwards k / v point to the
// sorted pairs and ka / va to the free buffers.
static void sort_pairs(Context& c, int64_t n, int endbit, uint64_t*& k, uint64_t*& ka, int32_t*& v, int32_t*& va,
                       const char* what) {
  hipcub::DoubleBuffer<uint64_t> dk(k, ka);
  hipcub::DoubleBuffer<int32_t> dv(v, va);
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, dk, dv, static_cast<int>(n), 0, endbit, c.stream);
  c.ad_tmp.ensure(std::max<size_t>(bytes, 1));
  hip_check(hipcub::DeviceRadixSort::SortPairs(c.ad_tmp.p, bytes, dk, dv, static_cast<int>(n), 0, endbit, c.stream),
            what);
  k = dk.Current();
  ka = dk.Alternate();
  v = dv.Current();
  va = dv.Alternate();
}

// Table of the current grid, per-instance thresholds and the norms (shared
// set-up of refine and coarsen); returns the kernel arguments.
static AdArgs adapt_args(Context& c, const std::vector<double>* thr, int32_t cap_new) {
  const int64_t E = c.E();
  if (c.ad_lvtab.cap == 0) to_device(c.ad_lvtab, c.hk.level_table(c.D), c.stream);
  uint64_t tcap = 16;
  while (tcap < 2 * static_cast<uint64_t>(cap_new)) tcap <<= 1;
  if (cap_new > 0) c.ad_hkey.ensure(tcap);
  c.ad_newkey.ensure(static_cast<size_t>(std::max<int32_t>(cap_new, 1)));
  c.ad_ctr.ensure(kCtr);
  if (thr) {  // the norm test (else the caller's flags decide)
    to_device(c.ad_thr, *thr, c.stream);
    c.indicator_sumsq_dev();
  }
  AdArgs a{};
  for (int d = 0; d < c.D; ++d) {
    a.slot[d] = c.g_slot[d].p;
    a.fbase[d] = c.g_fbase[d].p;
    a.fmap[d] = c.g_fmap[d].p;
  }
  a.inst = c.g_inst.p;
  a.sumsq = c.sumsq.p;
  a.thr = c.ad_thr.p;
  a.ukey = c.g_ukey.p;
  a.lvtab = c.ad_lvtab.p;
  a.hkey = c.ad_hkey.p;
  a.mask = tcap - 1;
  a.newkey = c.ad_newkey.p;
  a.ctr = c.ad_ctr.p;
  a.E = E;
  a.D = c.D;
  a.nb = c.nmax;
  a.nmax = c.nmax;
  a.cap_new = cap_new;
  hipLaunchKernelGGL(k_ad_fill32, dim3(1), dim3(kAB), 0, c.stream, c.ad_ctr.p, int64_t{kCtr}, 0);
  if (cap_new > 0)
    hipLaunchKernelGGL(k_ad_fill64, dim3(ablocks(static_cast<int64_t>(tcap))), dim3(kAB), 0, c.stream, c.ad_hkey.p,
                       static_cast<int64_t>(tcap), kFree);
  return a;
}

static OutArgs out_args(Context& c, int64_t N) {
  const size_t n = static_cast<size_t>(std::max<int64_t>(N, 1));
  OutArgs o{};
  for (int d = 0; d < c.D; ++d) {
    c.ad_slot[d].ensure(n);
    o.slot[d] = c.ad_slot[d].p;
  }
  c.ad_inst.ensure(n);
  c.ad_lsum.ensure(n);
  c.g_remap.ensure(n);
  c.ad_ukey.ensure(n);
  // read-back words: [B] first and [B] last element of every instance (-1:
  // none), [B * D] deepest levels, then per dimension fibers and largest fiber
  const int64_t nmeta = static_cast<int64_t>(c.B()) * (2 + c.D) + 2 * c.D;
  c.ad_meta.ensure(static_cast<size_t>(nmeta));
  hipLaunchKernelGGL(k_ad_fill32, dim3(ablocks(nmeta)), dim3(kAB), 0, c.stream, c.ad_meta.p, nmeta, 0);
  hipLaunchKernelGGL(k_ad_fill32, dim3(ablocks(2 * c.B())), dim3(kAB), 0, c.stream, c.ad_meta.p,
                     int64_t{2} * c.B(), -1);
  o.inst = c.ad_inst.p;
  o.lsum = c.ad_lsum.p;
  o.remap = c.g_remap.p;
  o.ukey = c.ad_ukey.p;
  o.B = c.B();
  return o;
}

static std::array<int32_t, 2> read_counters(Context& c) {
  std::array<int32_t, 2> h{0, 0};
  hip_check(hipMemcpyAsync(h.data(), c.ad_ctr.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream), "download");
  c.sync();
  return h;
}

// elements with zmark == 0 keep their order; the new table then gets its
// fiber tables (N survivors)
static void compact_survivors(Context& c, const AdArgs& a, int64_t N) {
  const int64_t oldE = a.E;
  c.ad_vals[0].ensure(static_cast<size_t>(oldE));
  c.ad_vals[1].ensure(static_cast<size_t>(oldE));
  hipLaunchKernelGGL(k_ad_keep, dim3(ablocks(oldE)), dim3(kAB), 0, c.stream, c.ad_zmark.p, c.ad_vals[0].p, oldE);
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, c.ad_vals[0].p, c.ad_vals[1].p, static_cast<int>(oldE),
                                         c.stream);
  c.ad_tmp.ensure(std::max<size_t>(bytes, 1));
  hip_check(hipcub::DeviceScan::ExclusiveSum(c.ad_tmp.p, bytes, c.ad_vals[0].p, c.ad_vals[1].p,
                                             static_cast<int>(oldE), c.stream),
            "scan (survivors)");
  const OutArgs o = out_args(c, N);
  hipLaunchKernelGGL(k_ad_compact, dim3(ablocks(oldE)), dim3(kAB), 0, c.stream, a, c.ad_vals[1].p, o);
  c.device_grid_finish(N, oldE);
}

int64_t Context::refine_device(const std::vector<double>& eps) { return refine_device_impl(&eps, nullptr); }

int64_t Context::refine_device_flags(const int32_t* flags) { return refine_device_impl(nullptr, flags); }

int64_t Context::refine_device_impl(const std::vector<double>* eps, const int32_t* flags) {
  const int64_t oldE = E();
  if (!device_adapt_fits(oldE)) return -1;
  // new-element list of E/4 (at least 2^16); a refinement that adds more
  // runs again with a list four times larger (nothing has changed yet)
  int32_t cap_new = static_cast<int32_t>(std::min<int64_t>(std::max<int64_t>(oldE / 4, 65536), kMaxNew));
  std::array<int32_t, 2> h{0, 0};
  for (;;) {
    AdArgs a = adapt_args(*this, eps, cap_new);
    a.flags = flags;
    hipLaunchKernelGGL(k_ad_children, dim3(ablocks(oldE)), dim3(kAB), 0, stream, a);
    hipLaunchKernelGGL(k_ad_snap, dim3(1), dim3(1), 0, stream, a.ctr, cap_new);
    hipLaunchKernelGGL(k_ad_close_first, dim3(std::min<unsigned>(ablocks(oldE), 1024)), dim3(kAB), 0, stream, a);
    hipLaunchKernelGGL(k_ad_close, dim3(1), dim3(kCloseThreads), 0, stream, a, D * nmax);
    h = read_counters(*this);
    const bool full = (h[1] & AF_FULL) || h[0] > cap_new;
    if (full && !(h[1] & AF_BAND) && cap_new < kMaxNew) {
      cap_new = static_cast<int32_t>(std::min<int64_t>(int64_t{cap_new} * 4, kMaxNew));
      continue;
    }
    if (full || (h[1] & (AF_BAND | AF_DUP | AF_OPEN)) || !device_adapt_fits(oldE + h[0])) {
      ++stat_dev_fallback;
      dev_fallback_flags |= h[1] | (full ? AF_FULL : 0);
      return -1;
    }
    break;
  }
  ++stat_dev_adapt;
  if (h[0] == 0) return 0;
  HostTimer tg(*this, HOST_GRID);
  const int64_t N = oldE + h[0];
  for (int i = 0; i < 2; ++i) {
    ad_keys[i].ensure(static_cast<size_t>(N));
    ad_vals[i].ensure(static_cast<size_t>(N));
  }
  UnionArgs ua{};
  for (int d = 0; d < D; ++d) ua.slot[d] = g_slot[d].p;
  ua.inst = g_inst.p;
  ua.newkey = ad_newkey.p;
  ua.lvtab = ad_lvtab.p;
  ua.keys = ad_keys[0].p;
  ua.vals = ad_vals[0].p;
  ua.ctr = ad_ctr.p;
  ua.E_old = oldE;
  ua.N = N;
  ua.D = D;
  ua.nb = nmax;
  hipLaunchKernelGGL(k_ad_union_keys, dim3(ablocks(N)), dim3(kAB), 0, stream, ua);
  uint64_t *sk = ad_keys[0].p, *ska = ad_keys[1].p;
  int32_t *sv = ad_vals[0].p, *sva = ad_vals[1].p;
  sort_pairs(*this, N, 32 + bits_for(B()), sk, ska, sv, sva, "sort (element keys)");
  const OutArgs o = out_args(*this, N);
  hipLaunchKernelGGL(k_ad_union_gather, dim3(ablocks(N)), dim3(kAB), 0, stream, ua, sk, sv, o);
  device_grid_finish(N, oldE);
  return h[0];
}

int64_t Context::coarsen_device(const std::vector<double>& eta) {
  const int64_t oldE = E();
  if (!device_adapt_fits(oldE)) return -1;
  AdArgs a = adapt_args(*this, &eta, 0);
  ad_nchild.ensure(static_cast<size_t>(oldE));
  ad_zmark.ensure(static_cast<size_t>(oldE));
  ad_dec.ensure(static_cast<size_t>(oldE));
  ad_cand.ensure(static_cast<size_t>(oldE));
  a.nchild = ad_nchild.p;
  a.zmark = ad_zmark.p;
  a.dec = ad_dec.p;
  a.cand = ad_cand.p;
  hipLaunchKernelGGL(k_ad_fill32, dim3(ablocks(oldE)), dim3(kAB), 0, stream, ad_nchild.p, oldE, 0);
  hipLaunchKernelGGL(k_ad_nchild, dim3(ablocks(oldE)), dim3(kAB), 0, stream, a);
  const bool small = oldE <= kSmallTable;
  if (small) {
    ad_vals[0].ensure(static_cast<size_t>(oldE));
    ad_vals[1].ensure(static_cast<size_t>(oldE));
  }
  hipLaunchKernelGGL(k_ad_below, dim3(ablocks(oldE)), dim3(kAB), 0, stream, a, small ? ad_vals[0].p : nullptr);
  // an element removed in pass t has a chain of t - 1 removed descendants
  // below it, so passes 1 .. D * NMAX remove everything and pass D * NMAX + 1
  // finds nothing
  const int passes = D * nmax + 1;
  if (small) {
    hipLaunchKernelGGL(k_ad_passes, dim3(1), dim3(kCloseThreads), 0, stream, a, passes, ad_vals[0].p, ad_vals[1].p);
  } else {
    const unsigned g = std::min<unsigned>(ablocks(oldE), 1024);
    for (int it = 1; it <= passes; ++it) {
      hipLaunchKernelGGL(k_ad_mark, dim3(g), dim3(kAB), 0, stream, a, it, it == passes ? 1 : 0);
      if (it < passes) hipLaunchKernelGGL(k_ad_unlink, dim3(g), dim3(kAB), 0, stream, a, it);
    }
  }
  const std::array<int32_t, 2> h = read_counters(*this);
  if (h[1] & (AF_BAND | AF_FULL | AF_DUP | AF_OPEN) || h[0] >= oldE) {
    ++stat_dev_fallback;
    dev_fallback_flags |= h[1];
    return -1;
  }
  ++stat_dev_adapt;
  if (h[0] == 0) return 0;
  HostTimer tg(*this, HOST_GRID);
  compact_survivors(*this, a, oldE - h[0]);
  return h[0];
}

int64_t Context::drop_instances_device(const std::vector<char>& drop) {
  const int64_t oldE = E();
  if (static_cast<int>(drop.size()) != B()) throw std::runtime_error("drop instances: one flag per instance expected");
  if (!device_adapt_fits(oldE)) return -1;
  HostTimer tg(*this, HOST_GRID);
  int64_t removed = 0;
  std::vector<int32_t> flags(drop.size());
  for (int i = 0; i < B(); ++i) {
    flags[i] = drop[i] ? 1 : 0;
    if (drop[i]) removed += off[i + 1] - off[i];
  }
  if (removed == 0) return 0;
  ad_run.ensure(flags.size());
  to_device(ad_run, flags, stream);
  AdArgs a{};
  for (int d = 0; d < D; ++d) a.slot[d] = g_slot[d].p;
  a.inst = g_inst.p;
  a.ukey = g_ukey.p;
  ad_zmark.ensure(static_cast<size_t>(oldE));
  a.zmark = ad_zmark.p;
  a.E = oldE;
  a.D = D;
  hipLaunchKernelGGL(k_ad_drop, dim3(ablocks(oldE)), dim3(kAB), 0, stream, a, ad_run.p);
  compact_survivors(*this, a, oldE - removed);  // (waits for the stream: the flag upload is done)
  return removed;
}

void Context::device_grid_finish(int64_t N, int64_t oldE) {
  // fiber tables of every dimension: sort by (instance, slots outside d),
  // runs of equal keys are the fibers
  const int B_ = B();
  int32_t* meta_lmax = ad_meta.p + 2 * static_cast<int64_t>(B_);
  int32_t* meta_nfib = meta_lmax + static_cast<int64_t>(B_) * D;
  int32_t* meta_maxfib = meta_nfib + D;
  hipLaunchKernelGGL(k_ad_bounds, dim3(ablocks(N)), dim3(kAB), 0, stream, ad_inst.p, N, ad_meta.p, ad_meta.p + B_);
  for (int i = 0; i < 2; ++i) {
    ad_keys[i].ensure(static_cast<size_t>(N));
    ad_vals[i].ensure(static_cast<size_t>(N));
  }
  ad_run.ensure(static_cast<size_t>(N));
  ad_runmax.ensure(static_cast<size_t>(N));
  ad_runlen.ensure(static_cast<size_t>(N));
  const int endbit = D * nmax + bits_for(B_);
  size_t cbytes = 0;
  (void)hipcub::DeviceScan::InclusiveSum(nullptr, cbytes, ad_run.p, ad_run.p, static_cast<int>(N), stream);
  for (int dd = 0; dd < D; ++dd) {
    FibArgs f{};
    for (int d = 0; d < D; ++d) f.slot[d] = ad_slot[d].p;
    f.inst = ad_inst.p;
    f.keys = ad_keys[0].p;
    f.vals = ad_vals[0].p;
    f.N = N;
    f.D = D;
    f.nb = nmax;
    f.dd = dd;
    hipLaunchKernelGGL(k_ad_fib_keys, dim3(ablocks(N)), dim3(kAB), 0, stream, f);
    uint64_t *sk = ad_keys[0].p, *ska = ad_keys[1].p;
    int32_t *sv = ad_vals[0].p, *sva = ad_vals[1].p;
    sort_pairs(*this, N, endbit, sk, ska, sv, sva, "sort (fiber keys)");
    // run starts into the free value buffer, then their inclusive sum
    hipLaunchKernelGGL(k_ad_run_flags, dim3(ablocks(N)), dim3(kAB), 0, stream, sk, sva, N, ad_runmax.p, ad_runlen.p);
    ad_tmp.ensure(std::max<size_t>(cbytes, 1));
    size_t b2 = cbytes;
    hip_check(hipcub::DeviceScan::InclusiveSum(ad_tmp.p, b2, sva, ad_run.p, static_cast<int>(N), stream),
              "scan (fiber runs)");
    ad_fbase[dd].ensure(static_cast<size_t>(N));
    ad_fmaxl[dd].ensure(static_cast<size_t>(N));
    hipLaunchKernelGGL(k_ad_runs, dim3(ablocks(N)), dim3(kAB), 0, stream, sv, ad_run.p, ad_slot[dd].p,
                       ad_runmax.p, ad_runlen.p, ad_fbase[dd].p, N, nmax, meta_nfib + dd);
    hipLaunchKernelGGL(k_ad_run_levels, dim3(ablocks(N)), dim3(kAB), 0, stream, sv, ad_run.p, ad_runmax.p,
                       ad_runlen.p, ad_fmaxl[dd].p, N, meta_maxfib + dd, ad_inst.p, meta_lmax, D, dd);
  }
  // counts per instance, deepest levels, fibers and largest fiber per dimension
  const int64_t nmeta = static_cast<int64_t>(B_) * (2 + D) + 2 * D;
  const size_t at_nfib = static_cast<size_t>(B_) * (2 + D), at_maxfib = at_nfib + D;
  std::vector<int32_t> meta(static_cast<size_t>(nmeta));
  std::array<int32_t, 2> h{0, 0};
  hip_check(hipMemcpyAsync(meta.data(), ad_meta.p, nmeta * sizeof(int32_t), hipMemcpyDeviceToHost, stream),
            "download");
  hip_check(hipMemcpyAsync(h.data(), ad_ctr.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, stream), "download");
  sync();
  if (h[1] & AF_DUP) throw std::runtime_error("grid: duplicate hash key");
  if (h[1] & AF_NEG) throw std::runtime_error("grid: negative hash key");
  for (int d = 0; d < D; ++d)
    if ((int64_t{meta[at_nfib + d]} << nmax) > (int64_t{1} << 31) - 1)
      throw std::runtime_error("fiber map exceeds int32 range");
  std::vector<int64_t> noff(B_ + 1, 0);
  for (int i = 0; i < B_; ++i) {
    const int32_t f = meta[i], l = meta[B_ + i];
    if (f >= 0 && f != noff[i]) throw std::logic_error("device refine/coarsen: instances out of order");
    noff[i + 1] = noff[i] + (f >= 0 ? l - f + 1 : 0);
  }
  if (noff[B_] != N) throw std::logic_error("device refine/coarsen: instance counts do not add up");
  // commit: fiber maps of the new table, then swap it in
  for (int d = 0; d < D; ++d) {
    const int64_t nfib = meta[at_nfib + d];
    const int64_t nmap = std::max<int64_t>(nfib << nmax, 1);
    g_fmap[d].ensure(static_cast<size_t>(nmap));
    const int64_t nwords = (nmap + 31) / 32;
    g_fbits[d].ensure(static_cast<size_t>(nwords + 1));
    hipLaunchKernelGGL(k_ad_clear_map, dim3(ablocks(std::max(nmap, nwords + 1))), dim3(kAB), 0, stream, g_fmap[d].p,
                       nmap, g_fbits[d].p, nwords + 1);
    fbits_words[d] = nwords;
    hipLaunchKernelGGL(k_ad_fmap, dim3(ablocks(N)), dim3(kAB), 0, stream, g_fmap[d].p, g_fbits[d].p, ad_fbase[d].p,
                       ad_slot[d].p, N);
    g_slot[d].swap(ad_slot[d]);
    g_fbase[d].swap(ad_fbase[d]);
    g_fmaxl[d].swap(ad_fmaxl[d]);
  }
  g_inst.swap(ad_inst);
  g_lsum.swap(ad_lsum);
  g_ukey.swap(ad_ukey);
  maxfib = 1;
  for (int d = 0; d < D; ++d) maxfib = std::max(maxfib, meta[at_maxfib + d]);
  inst_lmax.assign(meta.begin() + 2 * static_cast<int64_t>(B_), meta.begin() + at_nfib);
  off = noff;
  remap_field_dev(u, K, N, oldE);
  remap_field_dev(pred, K, N, oldE);
  ++grid_version;
  host_stale = true;
  ensure_fields();
  if (is_varcoef(flux_kind)) compute_vcoef();
}

void Context::ensure_host_grid() {
  if (!host_stale) return;
  const int64_t n = E();
  std::vector<std::vector<int32_t>> dsl(D, std::vector<int32_t>(static_cast<size_t>(n)));
  for (int d = 0; d < D; ++d)
    if (n > 0)
      hip_check(hipMemcpyAsync(dsl[d].data(), g_slot[d].p, n * sizeof(int32_t), hipMemcpyDeviceToHost, stream),
                "download");
  sync();
  std::vector<HostGrid> ng(B());
  parallel_for(B(), [&](int i) {
    std::vector<int32_t> s(static_cast<size_t>(off[i + 1] - off[i]) * D);
    for (int64_t e = off[i]; e < off[i + 1]; ++e)
      for (int d = 0; d < D; ++d) s[(e - off[i]) * D + d] = dsl[d][e];
    ng[i].build(D, nmax, s, hk);
  });
  grids = std::move(ng);
  host_union();
  host_stale = false;
}

bool Context::set_grids_fresh(const std::vector<std::vector<int32_t>>& slots) {
  const int nb = static_cast<int>(slots.size());
  int64_t N = 0;
  for (const std::vector<int32_t>& v : slots) N += static_cast<int64_t>(v.size()) / D;
  if (nb != B() || !up_host.empty() || !device_adapt_fits(N)) {
    set_grids(slots, false);
    return false;
  }
  // element keys of the new table, packed on the host
  std::vector<uint64_t> keys(static_cast<size_t>(N));
  int64_t q = 0;
  for (int i = 0; i < nb; ++i)
    for (size_t e = 0; e < slots[i].size() / D; ++e, ++q) {
      uint64_t k = static_cast<uint64_t>(i) << (D * nmax);
      for (int d = 0; d < D; ++d) {
        const int32_t sl = slots[i][e * D + d];
        if (sl < 0 || sl >= (1 << nmax)) throw std::runtime_error("grid: slot out of range");
        k |= static_cast<uint64_t>(sl) << (nmax * d);
      }
      keys[q] = k;
    }
  HostTimer tg(*this, HOST_GRID);
  if (ad_lvtab.cap == 0) to_device(ad_lvtab, hk.level_table(D), stream);
  ad_ctr.ensure(kCtr);
  hipLaunchKernelGGL(k_ad_fill32, dim3(1), dim3(kAB), 0, stream, ad_ctr.p, int64_t{kCtr}, 0);
  to_device(ad_newkey, keys, stream);
  for (int i = 0; i < 2; ++i) {
    ad_keys[i].ensure(static_cast<size_t>(N));
    ad_vals[i].ensure(static_cast<size_t>(N));
  }
  UnionArgs ua{};
  for (int d = 0; d < D; ++d) ua.slot[d] = g_slot[d].p;
  ua.inst = g_inst.p;
  ua.newkey = ad_newkey.p;
  ua.lvtab = ad_lvtab.p;
  ua.keys = ad_keys[0].p;
  ua.vals = ad_vals[0].p;
  ua.ctr = ad_ctr.p;
  ua.E_old = 0;
  ua.N = N;
  ua.D = D;
  ua.nb = nmax;
  hipLaunchKernelGGL(k_ad_union_keys, dim3(ablocks(N)), dim3(kAB), 0, stream, ua);
  uint64_t *sk = ad_keys[0].p, *ska = ad_keys[1].p;
  int32_t *sv = ad_vals[0].p, *sva = ad_vals[1].p;
  sort_pairs(*this, N, 32 + bits_for(B()), sk, ska, sv, sva, "sort (element keys)");
  const OutArgs o = out_args(*this, N);
  hipLaunchKernelGGL(k_ad_union_gather, dim3(ablocks(N)), dim3(kAB), 0, stream, ua, sk, sv, o);
  // (every element new: u and the predict copy come back zero)
  device_grid_finish(N, E());
  new_add.assign(static_cast<size_t>(N), 1);
  for (DevBuf<double>* b : {&tn, &rhs})
    hipLaunchKernelGGL(k_ad_fill64, dim3(ablocks(N * K)), dim3(kAB), 0, stream, reinterpret_cast<uint64_t*>(b->p),
                       N * K, uint64_t{0});
  return true;
}

void Context::norm_flags_dev(const std::vector<double>& thr, const int32_t* mask, int32_t* flags) {
  const int64_t n = E();
  AdArgs a = adapt_args(*this, &thr, 0);
  ad_band.ensure(static_cast<size_t>(std::max<int64_t>(n, 1)));
  hipLaunchKernelGGL(k_ad_norm_flags, dim3(ablocks(n)), dim3(kAB), 0, stream, a, mask, flags, ad_band.p);
  int32_t nband = 0;
  hip_check(hipMemcpyAsync(&nband, ad_ctr.p + 2, sizeof(int32_t), hipMemcpyDeviceToHost, stream), "download");
  sync();
  if (nband == 0) return;
  // norms within a few ulps of the threshold: the reference's pow decides
  std::vector<int32_t> ids(static_cast<size_t>(nband)), vals(static_cast<size_t>(nband));
  std::vector<double> s(static_cast<size_t>(n));
  hip_check(hipMemcpyAsync(ids.data(), ad_band.p, nband * sizeof(int32_t), hipMemcpyDeviceToHost, stream), "download");
  hip_check(hipMemcpyAsync(s.data(), sumsq.p, n * sizeof(double), hipMemcpyDeviceToHost, stream), "download");
  sync();
  for (int32_t q = 0; q < nband; ++q) {
    const int64_t e = ids[q];
    const int i = static_cast<int>(std::upper_bound(off.begin(), off.end(), e) - off.begin()) - 1;
    vals[q] = threshold_norm(s[e], thr[i]) > thr[i] ? 1 : 0;
  }
  DevBuf<int32_t> dids, dvals;
  to_device(dids, ids, stream);
  to_device(dvals, vals, stream);
  hipLaunchKernelGGL(k_ad_set_flags, dim3(ablocks(nband)), dim3(kAB), 0, stream, dids.p, dvals.p, nband, flags);
  sync();  // (the id and value buffers are freed on return)
}

int64_t Context::proposer_step_device(const std::function<void(int, int, double)>& stage,
                                      const std::vector<double>& eps) {
  const int64_t E0 = E();
  // leaves of the grid at the start of the step (the judge's candidates)
  for (DevBuf<int32_t>& b : ad_leaf) b.ensure(static_cast<size_t>(E0));
  ad_flag.ensure(static_cast<size_t>(E0));
  {
    AdArgs a = adapt_args(*this, nullptr, 0);
    hipLaunchKernelGGL(k_ad_leaf, dim3(ablocks(E0)), dim3(kAB), 0, stream, a, ad_leaf[0].p);
  }
  // proposals: leaves with a probability above the threshold (learned), or
  // with a current norm above eps, get their children
  {
    HostTimer timer(*this, HOST_ADAPT);
    if (refine_mode == 1) {
      prop_feats.ensure(static_cast<size_t>(std::max<int64_t>(E0, 1)) * kPropIn);
      prop_leaf.ensure(std::max<int64_t>(E0, 1));
      prop_prob.ensure(std::max<int64_t>(E0, 1));
      proposer_features(eps, prop_feats.p, prop_leaf.p);
      proposer_predict(prop_feats.p, E0, prop_prob.p);
      hipLaunchKernelGGL(k_ad_prob_flags, dim3(ablocks(E0)), dim3(kAB), 0, stream, prop_prob.p, prop_threshold,
                         ad_flag.p, E0);
    } else {
      norm_flags_dev(eps, nullptr, ad_flag.p);
    }
    const int64_t added = refine_device_flags(ad_flag.p);
    if (added < 0) throw std::runtime_error("device refinement of the proposals failed");
    stat_proposed += added;
    if (added > 0) {
      ad_leaf[1].ensure(static_cast<size_t>(E()));
      hipLaunchKernelGGL(k_ad_carry, dim3(ablocks(E())), dim3(kAB), 0, stream, g_remap.p, ad_leaf[0].p,
                         ad_leaf[1].p, E());
      ad_leaf[0].swap(ad_leaf[1]);
    }
  }
  rk_init();
  for (int st = 0; st < 3; ++st) stage(3, st, 0.);
  // the deferring judge: start-of-step leaves whose norm now exceeds eps get
  // their children after the step
  HostTimer timer(*this, HOST_ADAPT);
  ad_flag.ensure(static_cast<size_t>(E()));
  norm_flags_dev(eps, ad_leaf[0].p, ad_flag.p);
  const int64_t judged = refine_device_flags(ad_flag.p);
  if (judged < 0) throw std::runtime_error("device refinement of the judge failed");
  stat_deferred += judged;
  return 0;
}

std::vector<std::vector<int32_t>> Context::all_instance_slots() {
  std::vector<std::vector<int32_t>> out(B());
  if (!host_stale) {
    for (int i = 0; i < B(); ++i) out[i] = grids[i].slot;
    return out;
  }
  const int64_t n = E();
  std::vector<std::vector<int32_t>> col(D, std::vector<int32_t>(static_cast<size_t>(n)));
  for (int d = 0; d < D; ++d)
    if (n > 0)
      hip_check(hipMemcpyAsync(col[d].data(), g_slot[d].p, n * sizeof(int32_t), hipMemcpyDeviceToHost, stream),
                "download");
  sync();
  for (int i = 0; i < B(); ++i) {
    out[i].resize(static_cast<size_t>(off[i + 1] - off[i]) * D);
    for (int64_t e = off[i]; e < off[i + 1]; ++e)
      for (int d = 0; d < D; ++d) out[i][(e - off[i]) * D + d] = col[d][e];
  }
  return out;
}

std::vector<int32_t> Context::instance_slots(int i) {
  if (i < 0 || i >= B()) throw std::runtime_error("instance index out of range");
  if (!host_stale) return grids[i].slot;
  const int64_t n = off[i + 1] - off[i];
  std::vector<int32_t> s(static_cast<size_t>(n) * D), col(static_cast<size_t>(n));
  for (int d = 0; d < D; ++d) {
    if (n > 0)
      hip_check(hipMemcpyAsync(col.data(), g_slot[d].p + off[i], n * sizeof(int32_t), hipMemcpyDeviceToHost, stream),
                "download");
    sync();
    for (int64_t e = 0; e < n; ++e) s[e * D + d] = col[e];
  }
  return s;
}

int64_t Context::check_device_grid() {
  ensure_host_grid();
  const int64_t n = E();
  int64_t bad = 0;
  auto down = [&](const DevBuf<int32_t>& b, int64_t cnt) {
    std::vector<int32_t> h(static_cast<size_t>(cnt));
    if (cnt > 0)
      hip_check(hipMemcpyAsync(h.data(), b.p, cnt * sizeof(int32_t), hipMemcpyDeviceToHost, stream), "download");
    sync();
    return h;
  };
  std::vector<int64_t> fib0(B() + 1, 0);
  const std::vector<int32_t> inst = down(g_inst, n), lsum = down(g_lsum, n);
  std::vector<uint64_t> ukey(static_cast<size_t>(n));
  if (n > 0)
    hip_check(hipMemcpyAsync(ukey.data(), g_ukey.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost, stream), "download");
  sync();
  for (int64_t e = 0; e < n; ++e) {
    bad += inst[e] != uinst[e];
    bad += ukey[e] != ((static_cast<uint64_t>(uinst[e]) << 32) | static_cast<uint32_t>(uhash[e]));
    int ls = 0;
    for (int d = 0; d < D; ++d) ls += level(e, d);
    bad += lsum[e] != ls;
  }
  int mf = 1;
  for (int d = 0; d < D; ++d) {
    for (int i = 0; i < B(); ++i) fib0[i + 1] = fib0[i] + grids[i].nfib[d];
    const std::vector<int32_t> sl = down(g_slot[d], n), fb = down(g_fbase[d], n), fm = down(g_fmaxl[d], n);
    const int64_t nmap = std::max<int64_t>(fib0[B()] << nmax, 1);
    bad += fbits_words[d] != (nmap + 31) / 32;
    const std::vector<int32_t> map = down(g_fmap[d], nmap);
    std::vector<int32_t> want(static_cast<size_t>(nmap), -1);
    for (int i = 0; i < B(); ++i) {
      const HostGrid& g = grids[i];
      for (int64_t e = 0; e < g.E; ++e) {
        const int64_t u_ = off[i] + e;
        const int32_t f = g.fbase[d][e] + static_cast<int32_t>(fib0[i] << nmax);
        bad += sl[u_] != g.slot[e * D + d];
        bad += fb[u_] != f;
        bad += fm[u_] != g.fmaxl[d][e];
        want[static_cast<size_t>(f) + g.slot[e * D + d]] = static_cast<int32_t>(u_);
      }
      mf = std::max(mf, g.maxfib[d]);
    }
    for (int64_t q = 0; q < nmap; ++q) bad += map[q] != want[q];
  }
  bad += maxfib != mf;
  return bad;
}

}  // namespace adg
