// Device code of the fiber-tile kernels, shared by their own launches
// (fused_tile.hip) and the persistent stage kernel (fused_coop.hip).
//
// A sweep along dimension d couples only the elements of one d-fiber (the
// elements that share their cells in every other dimension).  The elements of
// every dimension are kept in fiber order (perm: sorted by fiber, then by
// slot), and consecutive whole fibers are grouped into tiles.  A workgroup
// takes one (term, tile): it stages the term's input values of the tile's
// elements in LDS with loads of whole element rows, then every thread
// evaluates one target element for RG rest indices (the indices over the
// dimensions other than the swept one) -- the PT1 outputs of the swept
// dimension each -- by walking the target's related source cells in ascending
// slot order and reading their values from LDS.  The position of a present
// cell in fiber order is the rank of its occupancy bit (pre[word] + popcount
// of the lower bits), so no fiber-map gather is needed.  The table blocks of a
// (target, source) pair are read from global memory (L2).
//
// Every output sums its sources in ascending slot order, k inner, exactly as
// the per-element kernels and the CPU oracle do, so the values are
// bit-identical.  The last sweeps of the partial sums write each term's values
// to its own buffer; the combine steps add the terms of an output in the
// reference order (the terms of one output come from tiles of different
// dimensions).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kern.hpp"

namespace adg {

constexpr int kTileThreads = 256;
// LDS of a tile workgroup: metadata plus the staged inputs of a chunk of rest
// indices (the chunk shrinks for large tiles)
constexpr int kTileLds = 48 * 1024;
constexpr int kMaxRest = 128;  // rest indices of one element: Tin / size_dd <= 5^3

__device__ __forceinline__ int bit_rank(const uint32_t* bits, const int32_t* pre, int x) {
  const uint32_t w = bits[x >> 5];
  return pre[x >> 5] + __popc(w & ((1u << (x & 31)) - 1u));
}

// Present source slots of a target, ascending, with their positions in fiber
// order: f(slot, position).  Short runs (ancestors, flux neighbours) test
// their bits one by one; longer runs (descendants) walk the occupancy words,
// whose set bits have consecutive positions.
template <class F>
__device__ __forceinline__ void for_each_present_pos(int st, int maxl, bool flx, int lu, const uint32_t* bits,
                                                     const int32_t* pre, int fb, F&& f) {
  for_each_related_ranges(st, maxl, flx, lu, [&](int lo, int hi) {
    if (hi - lo <= 2) {
      for (int s = lo; s < hi; ++s) {
        const int x = fb + s;
        const uint32_t w = bits[x >> 5];
        const uint32_t b = 1u << (x & 31);
        if (w & b) f(s, pre[x >> 5] + __popc(w & (b - 1u)));
      }
      return;
    }
    const int b0 = fb + lo, b1 = fb + hi;  // bit range [b0, b1)
    for (int w = b0 >> 5; w <= (b1 - 1) >> 5; ++w) {
      const uint32_t full = bits[w];
      uint32_t m = full;
      const int wb = w << 5;
      if (wb < b0) m &= ~0u << (b0 - wb);
      if (wb + 32 > b1) m &= ~0u >> (wb + 32 - b1);
      if (!m) continue;
      int pos = pre[w] + __popc(full & ((1u << __builtin_ctz(m)) - 1u));
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        f(wb + b - fb, pos++);
      }
    }
  });
}

__device__ __forceinline__ void rest_offsets_t(int r, int dd, int D, const int* isz, const int* si, const int* so,
                                               int& bin, int& bout) {
  bin = 0;
  bout = 0;
  for (int d = D - 1; d >= 0; --d) {
    if (d == dd) continue;
    const int q = r % isz[d];
    r /= isz[d];
    bin += q * si[d];
    bout += q * so[d];
  }
}

// Loads through the global address space: the descriptors hold generic
// pointers, whose flat loads would also count against the LDS wait counter.
template <class T>
__device__ __forceinline__ T gload(const T* p, int64_t o) {
  return ((const __attribute__((address_space(1))) T*)p)[o];
}
template <class T>
struct TileLoad;
template <>
struct TileLoad<double> {
  static __device__ __forceinline__ double in(const TermDev& t, int64_t o) { return gload(t.in, o); }
  static __device__ __forceinline__ SweepOps<double>::M mat(const TermDev& t, int64_t o) { return {gload(t.M, o)}; }
};
template <>
struct TileLoad<float> {
  static __device__ __forceinline__ float in(const TermDev& t, int64_t o) {
    return t.in32 ? gload(t.in32, o) : static_cast<float>(gload(t.in, o));
  }
  static __device__ __forceinline__ SweepOps<float>::M mat(const TermDev& t, int64_t o) {
    return {gload(t.Mh, o), gload(t.Ml, o)};
  }
};

struct TileSet {
  const int32_t* tbeg[kMaxDimK];
  const int32_t* tend[kMaxDimK];
};

// Metadata bytes of a tile workgroup with at most max_tile elements.
__host__ __device__ inline int tile_meta_bytes(int max_tile) { return ((max_tile * 3 * 4 + 15) / 16) * 16; }

// The related sources of target position p (slot st, walk depth maxl) with
// their positions, ascending: f(source slot, local position).
template <class F>
__device__ __forceinline__ void tile_sources(const TermDev& t, int st, int maxl, int fb, int p0, int n, F&& f) {
  for_each_present_pos(st, maxl, t.flx != 0, t.lu, t.fbits, t.pre, fb, [&](int s, int pos) {
    const int li = pos - p0;
    if (static_cast<unsigned>(li) < static_cast<unsigned>(n)) f(s, li);
  });
}

// A thread evaluates one target for RG consecutive rest indices of the
// staged chunk (RG = 1: one rest index per thread), so one walk over the
// fiber and one read of each table block serve RG rest indices.
// Work item vb (= task * ntile + tile) of a tile launch, evaluated by the
// whole workgroup; smem: lds_bytes of dynamic LDS, s_bin / s_bout: kMaxRest
// ints of LDS each.  Returns early (the whole workgroup) for an empty tile.
template <int PF1, int PT1, class T, int RG>
__device__ __forceinline__ void tile_item(const TaskDev* tasks, int64_t ntile, const TileSet& ts, int D, int cap,
                                          int max_tile, int lds_bytes, unsigned char* smem, int* s_bin, int* s_bout,
                                          int64_t vb) {
  using Ops = SweepOps<T>;
  const int64_t j = vb / ntile;
  const int64_t tile = vb - j * ntile;
  const TaskDev& tk = tasks[j];
  const TermDev& t = tk.t;
  const int dd = t.dd;
  const int p0 = ts.tbeg[dd][tile], p1 = ts.tend[dd][tile];
  if (p1 <= p0) return;  // no fiber starts in this tile's window
  const int n = p1 - p0;
  int32_t* s_elem = reinterpret_cast<int32_t*>(smem);
  int32_t* s_fb = s_elem + max_tile;
  int32_t* s_sm = s_fb + max_tile;  // slot | (walk depth + 1) << 16; depth 0: above the level-sum cap
  T* X = reinterpret_cast<T*>(smem + tile_meta_bytes(max_tile));
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int i = tid; i < n; i += nth) {
    const int e = t.perm[p0 + i];
    const int st = t.slot[e];
    int maxl = t.fmaxl[e];
    if (cap >= 0) {  // coarse grid (FastMultiplyLU::transform_1D_coarse_grid)
      const int ls = t.lsum[e];
      maxl = ls > cap ? -1 : min(maxl, cap - ls + slot_level(st));
    }
    s_elem[i] = e;
    s_fb[i] = t.fbase[e];
    s_sm[i] = st | ((maxl + 1) << 16);
  }
  int so[kMaxDimK];
  so[D - 1] = 1;
  for (int d = D - 2; d >= 0; --d) so[d] = so[d + 1] * tk.osz[d + 1];
  const int sod = so[dd];
  const int sdd = t.sin[dd];
  const int R = t.Tin / t.isz[dd];
  // rest indices per chunk, from this tile's own element count (n <= max_tile,
  // so at least the chunk the host checked for a tile of max_tile elements)
  int rc = min(R, (lds_bytes - tile_meta_bytes(max_tile)) / (n * PF1 * static_cast<int>(sizeof(T))));
  rc = min(rc, kMaxRest / PF1);  // a staged row has at most kMaxRest <= blockDim values
  const int64_t Mrow = static_cast<int64_t>(t.ns) * PF1 * PT1;
  T* out = Ops::out(tk);
  for (int r0 = 0; r0 < R; r0 += rc) {
    const int nr = min(rc, R - r0);
    __syncthreads();  // metadata written / previous chunk's reads done
    if (tid < nr) {
      int bi, bo;
      rest_offsets_t(r0 + tid, dd, D, t.isz, t.sin, so, bi, bo);
      s_bin[tid] = bi;
      s_bout[tid] = bo;
    }
    __syncthreads();
    {
      // staging: thread (q, c) copies value c of the rows q, q + step, ...
      // (per = nr * PF1 <= kMaxRest values of a row, step = rows per pass)
      const int per = nr * PF1;
      const int step = nth / per, q0 = tid / per, c = tid - q0 * per;
      if (q0 < step) {
        const int r = c / PF1, k = c - r * PF1;
        const int off = s_bin[r] + k * sdd;
        for (int p = q0; p < n; p += step)
          X[p * per + c] = TileLoad<T>::in(t, static_cast<int64_t>(s_elem[p]) * t.Tin + off);
      }
    }
    __syncthreads();
    // thread (q, g): rest indices g*RG .. g*RG + RG - 1 of the targets q, q + step, ...
    const int ng = (nr + RG - 1) / RG;
    const int step = nth / ng, q0 = tid / ng, g = tid - q0 * ng;
    const int rb = g * RG, nrg = min(RG, nr - rb);
    if (q0 < step) {
      for (int p = q0; p < n; p += step) {
        const int sm = s_sm[p];
        const int st = sm & 0xffff, maxl = (sm >> 16) - 1;
        const int e = s_elem[p];
        T acc[RG][PT1];
#pragma unroll
        for (int r = 0; r < RG; ++r)
#pragma unroll
          for (int i = 0; i < PT1; ++i) acc[r][i] = T(0);
        if (maxl >= 0) {
          const int64_t Mst = static_cast<int64_t>(st) * Mrow;
          tile_sources(t, st, maxl, s_fb[p], p0, n, [&](int s, int li) {
            const int64_t Mr = Mst + static_cast<int64_t>(s) * PF1 * PT1;
            typename Ops::M m[PF1][PT1];
#pragma unroll
            for (int k = 0; k < PF1; ++k)
#pragma unroll
              for (int i = 0; i < PT1; ++i) m[k][i] = TileLoad<T>::mat(t, Mr + k * PT1 + i);
            const T* x = X + (li * nr + rb) * PF1;
#pragma unroll
            for (int r = 0; r < RG; ++r) {
              if (r >= nrg) break;
              T xv[PF1];
#pragma unroll
              for (int k = 0; k < PF1; ++k) xv[k] = x[r * PF1 + k];
#pragma unroll
              for (int i = 0; i < PT1; ++i)
#pragma unroll
                for (int k = 0; k < PF1; ++k) Ops::madd(acc[r][i], xv[k], m[k][i]);
            }
          });
        }
        const T cf = T(term_coef(t, e));
        T* o = out + static_cast<int64_t>(e) * tk.Tout;
#pragma unroll
        for (int r = 0; r < RG; ++r) {
          if (r >= nrg) break;
#pragma unroll
          for (int i = 0; i < PT1; ++i) o[s_bout[rb + r] + i * sod] = maxl >= 0 ? acc[r][i] * cf : T(0);
        }
      }
    }
  }
}

// up[i] = sum over the nterm term values (term j at tv + j * ts) in order,
// then the point flux
template <class T>
__device__ __forceinline__ void up_combine_point(const T* tv, int nterm, int64_t n, int64_t ts, T* up, T* fp,
                                                 const double* vcoef, int kind, double fs, int ncomp, int64_t i) {
  double u = 0.;
  for (int j = 0; j < nterm; ++j) u += static_cast<double>(tv[j * ts + i]);
  const T uu = T(u);
  up[i] = uu;
  flux_point<T>(kind, uu, fs, vcoef, fp, i, n, ncomp);
}

// rhs[i] = sum over the term values and then the penalty values in order,
// then the RK stage update into u_next; non-finite values raise ERR_NONFINITE
template <class T>
__device__ __forceinline__ void rhs_combine_point(const T* tv, const double* pv, int nterm, int npen, int64_t n,
                                                  int K, double* rhs, const double* u, const double* tn,
                                                  double* u_next, int scheme, int stage, double dt, const double* dtv,
                                                  const int32_t* inst, int32_t* err, int64_t i) {
  const int64_t ts = n * static_cast<int64_t>(sizeof(double) / sizeof(T));
  double r = 0.;
  for (int j = 0; j < nterm; ++j) r += static_cast<double>(tv[j * ts + i]);
  for (int j = 0; j < npen; ++j) r += pv[j * n + i];
  rhs[i] = r;
  const double dte = dtv ? dtv[inst[i / K]] : dt;
  const double un = rk_combine(scheme, stage, dte, tn[i], u[i], r);
  u_next[i] = un;
  if (!isfinite(un)) atomicOr(err, ERR_NONFINITE);
}

}  // namespace adg
