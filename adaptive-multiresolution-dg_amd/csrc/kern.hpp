// Device-side descriptors of the fused RK stage and the launchers of its
// kernels.  Shared by the context (device.hip), the per-element gather kernels
// (fused_elem.hip) and the fiber-tile kernels (fused_tile.hip).
//
// A "term" is one 1D sweep (FastMultiplyLU::transform_1D,
// source/FastMultiplyLU.cpp:436-512) of one partial sum: out = coef * sum over
// the related source cells of the target's fiber, in ascending source-slot
// order, of in[src][.. k ..] * M(src slot, k; target slot, i).  A "task" is a
// term written to an output buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fiber.hpp"

namespace adg {

constexpr int kMaxDimK = 4;

struct TermDev {
  const double* in;
  const double* M;
  // FP32 precision mode: input in32 (or in, rounded on load, when in32 is
  // null) and the table split as M = Mh + Ml
  const float* in32;
  const float* Mh;
  const float* Ml;
  const int32_t* slot;
  const int32_t* fbase;
  const int32_t* fmaxl;
  const int32_t* fmap;
  const uint32_t* fbits;
  const int32_t* lsum;
  // fiber order of the swept dimension (fiber-tile kernels): elements sorted
  // by (fiber, slot), and the exclusive prefix popcounts of the occupancy
  // words (the position of a present cell in that order is pre[w] plus the
  // popcount of the lower bits of word w)
  const int32_t* perm;
  const int32_t* pre;
  // per-instance coefficient (replaces coef when non-null), indexed by inst[e]
  const double* coef_inst;
  const int32_t* inst;
  int ns, pt1, dd, Tin, flx, lu;
  int isz[kMaxDimK];  // sizes of the input tensor
  int sin[kMaxDimK];  // strides of the input tensor
  double coef;
};

struct TaskDev {  // a term written to an output buffer
  TermDev t;
  double* out;
  float* out32;  // FP32 precision mode
  int Tout;
  int osz[kMaxDimK];
  int R;               // rest indices per element (Tin / input size in the swept dimension)
  int64_t begin_v[4];  // first work item of this task in the phase, per per-element mapping
};

// Arithmetic of a sweep in the storage type T: FP64 (x * M added in order)
// or FP32 with the 1D table split as M = Mh + Ml in FP32, both parts applied
// by fused multiply-adds, the small part first.
template <class T>
struct SweepOps;
template <>
struct SweepOps<double> {
  struct M { double m; };
  static __device__ __forceinline__ double in(const TermDev& t, int64_t o) { return t.in[o]; }
  static __device__ __forceinline__ M mat(const TermDev& t, int64_t o) { return {t.M[o]}; }
  static __device__ __forceinline__ M zero() { return {0.}; }
  static __device__ __forceinline__ void madd(double& acc, double x, M m) { acc += x * m.m; }
  static __device__ __forceinline__ double* out(const TaskDev& tk) { return tk.out; }
};
template <>
struct SweepOps<float> {
  struct M { float h, l; };
  static __device__ __forceinline__ float in(const TermDev& t, int64_t o) {
    return t.in32 ? t.in32[o] : static_cast<float>(t.in[o]);
  }
  static __device__ __forceinline__ M mat(const TermDev& t, int64_t o) { return {t.Mh[o], t.Ml[o]}; }
  static __device__ __forceinline__ M zero() { return {0.f, 0.f}; }
  static __device__ __forceinline__ void madd(float& acc, float x, M m) {
    acc = fmaf(x, m.l, acc);
    acc = fmaf(x, m.h, acc);
  }
  static __device__ __forceinline__ float* out(const TaskDev& tk) { return tk.out32; }
};

// Term coefficient of target element e.
__device__ __forceinline__ double term_coef(const TermDev& t, int64_t e) {
  return t.coef_inst ? t.coef_inst[t.inst[e]] : t.coef;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// The gather kernels renumber their blocks so that every XCD works on one
// contiguous range of elements (for a batch: a contiguous range of whole
// instances), whose fiber neighbours then stay in that XCD's L2.  The
// renumbering is a bijection of block ids and changes no arithmetic.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned b = blockIdx.x, n = gridDim.x;
  const unsigned x = b % kXcds, i = b / kXcds, per = n / kXcds, extra = n % kXcds;
  return x * per + min(x, extra) + i;
}

// One explicit RK stage update from u_n (t), the current stage value (uu) and
// the right-hand side (rr): forward Euler (source/ODESolver.cpp:209-213),
// RK2 midpoint (:247-259) and SSP-RK3 (:285-301).
__device__ __forceinline__ double rk_combine(int scheme, int stage, double dt, double t, double uu, double rr) {
  // dt = 0 freezes an instance of a batch (its u and u_n both hold the
  // accepted solution while other instances redo their step)
  if (dt == 0.) return stage == 0 ? t : uu;
  if (scheme == 2) return stage == 0 ? t + 0.5 * dt * rr : t + dt * rr;
  if (scheme == 0 || stage == 0) return t + dt * rr;
  if (stage == 1) return 3. / 4. * t + 1. / 4. * (uu + dt * rr);
  return 1. / 3. * t + 2. / 3. * (uu + dt * rr);
}

// Point flux of the fused eval_up kernels (kind -1: none, the Hermite flux runs after).
// The values of kinds 0 and 1 are formed first and stored at fixed addresses
// (component 0 at fp[gid], component 1 at fp[n + gid]): when the branches
// stored themselves, the compiler merged the kind-0 store with the second
// kind-1 store into one store whose address came out as fp + n for both in
// the persistent stage kernel.
template <class T>
__device__ __forceinline__ void flux_point(int kind, T u, double s, const double* vcoef, T* fp, int64_t gid, int64_t n,
                                           int ncomp) {
  if (kind >= 2) {  // variable coefficient (scaled or not)
    for (int c = 0; c < ncomp; ++c) fp[c * n + gid] = T(vcoef[c * n + gid]) * u;
    return;
  }
  if (kind < 0) return;
  T f0, f1 = T(0);
  if (kind == 0) {
    f0 = u * u * T(0.5);
  } else {
    f0 = sin(u) * T(s);
    f1 = cos(u) * T(s);
  }
  fp[gid] = f0;
  if (kind == 1) fp[n + gid] = f1;
}

// Error flags of the device (bit values of the context's flag word).
// ERR_COOP: a persistent stage kernel found itself launched without the
// cooperative guarantee (it exits at once, leaving the stage undone).
enum : int32_t { ERR_ANCESTOR = 1, ERR_NONFINITE = 2, ERR_COOP = 4 };

// ---- per-element gather kernels (fused_elem.hip): thread mappings 0..3 ----
// v: 0 thread per output, 1 thread per (element, rest index), 2 thread per
// (element, 4 rest indices), 3 thread per (element, 2 rest indices); total
// counts the work items of that mapping.
template <class T>
void elem_phase(int pf1, int pt1, const TaskDev* tasks, int ntask, int64_t total, int D, int cap, int v, hipStream_t s);
template <class T>
void elem_up_final(int v, int k1, int m1, const TermDev* terms, int nterm, int64_t E, int P, int D, T* up, T* fp,
                   const double* vcoef, int kind, double fs, int ncomp, int cap, hipStream_t s);
template <class T>
void elem_rhs_final(int v, int m1, int k1, const TermDev* terms, int nterm, const TermDev* pens, int npen, int64_t E,
                    int K, int D, double* rhs, const double* u, const double* tn, double* u_next, int scheme, int stage,
                    double dt, const double* dtv, const int32_t* inst, int cap, int32_t* err, hipStream_t s);
// rest indices per thread of the per-element mappings 2 and 3
constexpr int kRC = 4;

// The fiber-tile kernels are declared in tile.hpp.

}  // namespace adg
