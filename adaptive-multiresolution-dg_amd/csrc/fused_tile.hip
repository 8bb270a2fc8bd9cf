// Launches of the fiber-tile kernels of the fused RK stage (device code in
// tile_dev.hpp) and the construction of the fiber order and the tiles.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <stdexcept>

#include "kern.hpp"
#include "tile_dev.hpp"
#include "tile.hpp"

namespace adg {
namespace {

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// LAST: the launch of the last sweeps (term values) rather than a phase of
// first sweeps; the code is the same, the separate instance tells the two
// apart in kernel traces.
template <int PF1, int PT1, class T, int RG, bool LAST>
__global__ __launch_bounds__(kTileThreads) void k_tile(const TaskDev* tasks, int64_t ntile, TileSet ts, int D,
                                                       int cap, int max_tile, int lds_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_bin[kMaxRest], s_bout[kMaxRest];
  tile_item<PF1, PT1, T, RG>(tasks, ntile, ts, D, cap, max_tile, lds_bytes, smem, s_bin, s_bout, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_word_popc(const uint32_t* bits, int64_t nwords, int32_t* cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nwords) cnt[i] = __popc(bits[i]);
  else if (i == nwords) cnt[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_fill_i32t(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// perm[rank(fbase + slot)] = e; the first element of every fiber extends the
// range of the tile its position falls in.
__global__ __launch_bounds__(kBlock) void k_fiber_order(const uint32_t* bits, const int32_t* pre,
                                                        const int32_t* fbase, const int32_t* slot, int64_t E, int ns,
                                                        int te, int32_t* perm, int32_t* tbeg, int32_t* tend) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int fb = fbase[e];
  const int pos = bit_rank(bits, pre, fb + slot[e]);
  perm[pos] = static_cast<int32_t>(e);
  if (pos == bit_rank(bits, pre, fb)) {
    const int t = pos / te;
    atomicMin(&tbeg[t], pos);
    atomicMax(&tend[t], bit_rank(bits, pre, fb + ns));
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_up_combine(const T* tv, int nterm, int64_t n, int64_t ts, T* up, T* fp,
                                                       const double* vcoef, int kind, double fs, int ncomp) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) up_combine_point<T>(tv, nterm, n, ts, up, fp, vcoef, kind, fs, ncomp, i);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_rhs_combine(const T* tv, const double* pv, int nterm, int npen, int64_t E,
                                                        int K, double* rhs, const double* u, const double* tn,
                                                        double* u_next, int scheme, int stage, double dt,
                                                        const double* dtv, const int32_t* inst, int32_t* err) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < E * K)
    rhs_combine_point<T>(tv, pv, nterm, npen, E * K, K, rhs, u, tn, u_next, scheme, stage, dt, dtv, inst, err, i);
}

template <int PF1, int PT1, class T, bool LAST>
void tile_sweep_t(const TaskDev* tasks, int ntask, int64_t ntile, const TileSet& ts, int max_tile, int D, int cap,
                  int lds, bool allr, hipStream_t s) {
  const int64_t nb = ntile * ntask;
  if (nb > INT_MAX) throw std::runtime_error("tile sweep: grid too large");
  if (allr)
    hipLaunchKernelGGL((k_tile<PF1, PT1, T, 4, LAST>), dim3(static_cast<unsigned>(nb)), dim3(kTileThreads), lds, s,
                       tasks, ntile, ts, D, cap, max_tile, lds);
  else
    hipLaunchKernelGGL((k_tile<PF1, PT1, T, 1, LAST>), dim3(static_cast<unsigned>(nb)), dim3(kTileThreads), lds, s,
                       tasks, ntile, ts, D, cap, max_tile, lds);
}

}  // namespace

size_t tile_scan_bytes(int64_t nwords) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const int32_t*>(nullptr),
                                   static_cast<int32_t*>(nullptr), static_cast<int>(nwords + 1));
  return bytes;
}

void tile_prefix(const uint32_t* bits, int64_t nwords, int32_t* cnt, int32_t* pre, void* scan_tmp,
                 size_t scan_tmp_bytes, hipStream_t s) {
  if (nwords + 1 > INT_MAX) throw std::runtime_error("tile build: fiber map too large");
  hipLaunchKernelGGL(k_word_popc, dim3(nblocks(nwords + 1)), dim3(kBlock), 0, s, bits, nwords, cnt);
  size_t bytes = scan_tmp_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(scan_tmp, bytes, cnt, pre, static_cast<int>(nwords + 1), s) != hipSuccess)
    throw std::runtime_error("tile build: prefix scan failed");
}

void tile_order(const uint32_t* bits, const int32_t* pre, const int32_t* fbase, const int32_t* slot, int64_t E, int ns,
                int te, int64_t ntile, int32_t* perm, int32_t* tbeg, int32_t* tend, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32t, dim3(nblocks(ntile)), dim3(kBlock), 0, s, tbeg, ntile, INT_MAX);
  hipLaunchKernelGGL(k_fill_i32t, dim3(nblocks(ntile)), dim3(kBlock), 0, s, tend, ntile, -1);
  if (E > 0)
    hipLaunchKernelGGL(k_fiber_order, dim3(nblocks(E)), dim3(kBlock), 0, s, bits, pre, fbase, slot, E, ns, te, perm,
                       tbeg, tend);
}

int tile_rest_chunk(int pf1, int R, int max_tile, int tbytes) {
  const int avail = kTileLds - tile_meta_bytes(max_tile);
  const int per = max_tile * pf1 * tbytes;
  // a staged row holds at most kMaxRest values of the chunk, one per thread
  if (R > kMaxRest || avail < per || pf1 > kMaxRest) return 0;
  return std::min(std::min(R, avail / per), kMaxRest / pf1);
}

template <class T>
void tile_sweep(int pf1, int pt1, const TaskDev* tasks, int ntask, int64_t ntile, int max_tile, const int32_t* const* tbeg,
                const int32_t* const* tend, int D, int cap, bool allr, bool last, hipStream_t s) {
  TileSet ts{};
  for (int d = 0; d < D; ++d) {
    ts.tbeg[d] = tbeg[d];
    ts.tend[d] = tend[d];
  }
  const int lds = kTileLds;
#define ADG_TS(F, P)                                                                \
  if (pf1 == F && pt1 == P) {                                                       \
    if (last) tile_sweep_t<F, P, T, true>(tasks, ntask, ntile, ts, max_tile, D, cap, lds, allr, s);  \
    else tile_sweep_t<F, P, T, false>(tasks, ntask, ntile, ts, max_tile, D, cap, lds, allr, s);      \
    return;                                                                         \
  }
  ADG_TS(2, 2) ADG_TS(2, 3) ADG_TS(2, 4) ADG_TS(2, 5) ADG_TS(3, 2) ADG_TS(3, 3) ADG_TS(3, 4) ADG_TS(3, 5)
  ADG_TS(4, 2) ADG_TS(4, 3) ADG_TS(4, 4) ADG_TS(4, 5) ADG_TS(5, 2) ADG_TS(5, 3) ADG_TS(5, 4)
#undef ADG_TS
  throw std::runtime_error("tile sweep: unsupported basis sizes");
}

template <class T>
void up_combine(const T* tv, int nterm, int64_t n, T* up, T* fp, const double* vcoef, int kind, double fs, int ncomp,
                hipStream_t s) {
  const int64_t ts = n * static_cast<int64_t>(sizeof(double) / sizeof(T));
  if (n > 0)
    hipLaunchKernelGGL(k_up_combine<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, tv, nterm, n, ts, up, fp, vcoef, kind,
                       fs, ncomp);
}

template <class T>
void rhs_combine(const T* tv, const double* pv, int nterm, int npen, int64_t E, int K, double* rhs, const double* u,
                 const double* tn, double* u_next, int scheme, int stage, double dt, const double* dtv,
                 const int32_t* inst, int32_t* err, hipStream_t s) {
  if (E > 0)
    hipLaunchKernelGGL(k_rhs_combine<T>, dim3(nblocks(E * K)), dim3(kBlock), 0, s, tv, pv, nterm, npen, E, K, rhs, u,
                       tn, u_next, scheme, stage, dt, dtv, inst, err);
}

template void tile_sweep<double>(int, int, const TaskDev*, int, int64_t, int, const int32_t* const*,
                                 const int32_t* const*, int, int, bool, bool, hipStream_t);
template void tile_sweep<float>(int, int, const TaskDev*, int, int64_t, int, const int32_t* const*,
                                const int32_t* const*, int, int, bool, bool, hipStream_t);
template void up_combine<double>(const double*, int, int64_t, double*, double*, const double*, int, double, int,
                                 hipStream_t);
template void up_combine<float>(const float*, int, int64_t, float*, float*, const double*, int, double, int,
                                hipStream_t);
template void rhs_combine<double>(const double*, const double*, int, int, int64_t, int, double*, const double*,
                                  const double*, double*, int, int, double, const double*, const int32_t*, int32_t*,
                                  hipStream_t);
template void rhs_combine<float>(const float*, const double*, int, int, int64_t, int, double*, const double*,
                                 const double*, double*, int, int, double, const double*, const int32_t*, int32_t*,
                                 hipStream_t);

}  // namespace adg
