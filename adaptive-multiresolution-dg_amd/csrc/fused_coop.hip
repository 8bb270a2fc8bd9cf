// Persistent RK stage: one cooperative launch per stage.
//
// Every workgroup of the launch loops over the work items of a phase (tiles
// of whole fibers for the sweeps, points for the element-wise steps), then
// the grid waits at a barrier before the next phase reads what the phase
// wrote.  The phases and their arithmetic are those of the separate launches
// (tile_dev.hpp, surplus_dev.hpp), so the results are bit-identical.  The
// launch is cooperative: the runtime admits it only when every workgroup is
// resident, so the barriers cannot wait on a workgroup that never runs.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <stdexcept>

#include "coop.hpp"

namespace adg {
namespace {

namespace cg = cooperative_groups;

template <int K1, int M1, class T>
__global__ __launch_bounds__(kTileThreads) void k_stage_coop(CoopStage S) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_bin[kMaxRest], s_bout[kMaxRest];
  cg::grid_group grid = cg::this_grid();
  // every workgroup takes the same branch: a launch without the cooperative
  // guarantee (no grid barrier workspace) is reported instead of waiting
  if (!grid.is_valid()) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(S.err, ERR_COOP);
    return;
  }
  const int64_t nthreads = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t gtid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int p = 0; p < S.nphase; ++p) {
    const CoopPhase& ph = S.ph[p];
    if (ph.kind == COOP_TILE_UP || ph.kind == COOP_TILE_RHS) {
      const int64_t n = ph.items + ph.items2;
      for (int64_t vb = blockIdx.x; vb < n; vb += gridDim.x) {
        if (ph.kind == COOP_TILE_UP)
          tile_item<K1, M1, T, 4>(ph.tasks, S.ntile, S.ts, S.D, S.cap, S.max_tile, S.lds_bytes, smem, s_bin, s_bout,
                                  vb);
        else if (vb < ph.items)
          tile_item<M1, K1, T, 4>(ph.tasks, S.ntile, S.ts, S.D, S.cap, S.max_tile, S.lds_bytes, smem, s_bin, s_bout,
                                  vb);
        else
          tile_item<K1, K1, double, 4>(ph.tasks2, S.ntile, S.ts, S.D, S.cap, S.max_tile, S.lds_bytes, smem, s_bin,
                                       s_bout, vb - ph.items);
        __syncthreads();  // the next item rewrites the workgroup's LDS
      }
    } else if (ph.kind == COOP_UP_COMBINE) {
      const int64_t ts = S.nP * static_cast<int64_t>(sizeof(double) / sizeof(T));
      for (int64_t i = gtid; i < S.nP; i += nthreads)
        up_combine_point<T>(static_cast<const T*>(S.tv), S.n_up_terms, S.nP, ts, static_cast<T*>(S.up),
                            static_cast<T*>(S.fp), S.vcoef, S.up_kind, S.fs, S.ncomp, i);
    } else if (ph.kind == COOP_SURPLUS) {
      for (int64_t i = gtid; i < ph.items; i += nthreads) surplus_point<T>(S.sur[ph.dim], i);
    } else {
      for (int64_t i = gtid; i < S.nK; i += nthreads)
        rhs_combine_point<T>(static_cast<const T*>(S.tv), S.pv, S.n_rhs_terms, S.npen, S.nK, S.K, S.rhs, S.u, S.tn,
                             S.unext, S.scheme, S.stage, S.dt, S.dtv, S.inst, S.err, i);
    }
    grid.sync();
  }
}

template <int K1, int M1, class T>
void launch_coop(const CoopStage& st, int device, hipStream_t stream) {
  const void* fn = reinterpret_cast<const void*>(&k_stage_coop<K1, M1, T>);
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stage_coop<K1, M1, T>, kTileThreads, st.lds_bytes) !=
          hipSuccess ||
      per_cu < 1)
    throw std::runtime_error("persistent stage: no resident workgroup fits");
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu < 1)
    throw std::runtime_error("persistent stage: no compute units");
  CoopStage arg = st;
  void* args[] = {&arg};
  const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(static_cast<unsigned>(per_cu * ncu)), dim3(kTileThreads),
                                                  args, static_cast<unsigned>(st.lds_bytes), stream);
  if (e != hipSuccess) throw std::runtime_error(std::string("persistent stage launch: ") + hipGetErrorString(e));
}

}  // namespace

template <class T>
void coop_stage(int k1, int m1, const CoopStage& st, int device, hipStream_t stream) {
#define ADG_CS(K1, M1)                             \
  if (k1 == K1 && m1 == M1) {                      \
    launch_coop<K1, M1, T>(st, device, stream);    \
    return;                                        \
  }
  ADG_CS(2, 2) ADG_CS(2, 3) ADG_CS(2, 4) ADG_CS(2, 5) ADG_CS(3, 2) ADG_CS(3, 3) ADG_CS(3, 4) ADG_CS(3, 5)
  ADG_CS(4, 2) ADG_CS(4, 3) ADG_CS(4, 4) ADG_CS(4, 5)
#undef ADG_CS
  throw std::runtime_error("persistent stage: unsupported basis sizes");
}

template void coop_stage<double>(int, int, const CoopStage&, int, hipStream_t);
template void coop_stage<float>(int, int, const CoopStage&, int, hipStream_t);

}  // namespace adg
