// Persistent RK stage: one cooperative launch runs every phase of a fused
// stage (fused_coop.hip), the phases separated by grid-wide barriers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kern.hpp"
#include "surplus_dev.hpp"
#include "tile_dev.hpp"

namespace adg {

// One phase: tile work items (kinds 0..2: one workgroup per (task, tile)),
// or points of an element-wise step (kinds 3..5).  A tile phase may hold a
// second task list (the penalty sweeps beside the last right-hand-side sweeps).
enum : int {
  COOP_TILE_UP = 0,    // (k+1) -> (m+1) sweeps
  COOP_TILE_RHS = 1,   // (m+1) -> (k+1) sweeps (+ the penalties, (k+1) -> (k+1), FP64)
  COOP_UP_COMBINE = 3,
  COOP_SURPLUS = 4,
  COOP_RHS_COMBINE = 5
};

struct CoopPhase {
  int kind, dim;
  const TaskDev* tasks;
  const TaskDev* tasks2;  // penalties of a COOP_TILE_RHS phase (may be null)
  int64_t items, items2;  // work items of tasks / tasks2; points of an element-wise phase
};

constexpr int kCoopMaxPhases = 16;

struct CoopStage {
  int nphase;
  CoopPhase ph[kCoopMaxPhases];
  SurplusArgs sur[kMaxDimK];  // per dimension (FP64 or FP32 fields as the storage type)
  TileSet ts;
  int64_t ntile;
  int max_tile, lds_bytes, D, cap;
  // point values and flux
  const void* tv;  // term values of the last sweeps (storage type)
  void* up;
  void* fp;
  const double* vcoef;
  int up_kind, ncomp, n_up_terms;
  double fs;
  int64_t nP;
  // right-hand side and RK update
  const double* pv;
  int n_rhs_terms, npen, K;
  int64_t nK;
  double* rhs;
  const double* u;
  const double* tn;
  double* unext;
  int scheme, stage;
  double dt;
  const double* dtv;
  const int32_t* inst;
  int32_t* err;
};

static_assert(sizeof(CoopStage) <= 4096, "the stage descriptor is passed as a kernel argument");

// Launches the stage cooperatively on `stream` with as many workgroups as
// fit on the device at once (every workgroup loops over the work items of
// each phase).  Throws when the basis sizes have no instantiation.
template <class T>
void coop_stage(int k1, int m1, const CoopStage& st, int device, hipStream_t stream);

}  // namespace adg
