// Device code of one dimension of the hierarchical-surplus transform, shared
// by its own kernel (device.hip) and the persistent stage kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kern.hpp"

namespace adg {

// One dimension of the hierarchical-surplus transform
// (LagrInterpolation::eval_fp_to_coe_D_Lag, source/Interplation.cpp:1279-1400).
struct SurplusArgs {
  const double* in;
  double* out;
  const float* in32;  // FP32 storage (precision mode): in32 / out32 and the split weights
  float* out32;
  const float* wh;
  const float* wl;
  const int32_t* slot;
  const int32_t* fbase;
  const int32_t* fmap;
  const int32_t* anc_slot;
  const int32_t* anc_point;
  const double* wt;
  int32_t* err;
  int64_t E, comp_stride;
  int P, p1, stride_d, ncomp;
};

template <class T>
__device__ __forceinline__ const T* surplus_in(const SurplusArgs& a);
template <>
__device__ __forceinline__ const double* surplus_in<double>(const SurplusArgs& a) { return a.in; }
template <>
__device__ __forceinline__ const float* surplus_in<float>(const SurplusArgs& a) { return a.in32; }

// FP64: coe += wt * anc.  FP32: the weight is split, wt = wh + wl, and both
// parts are applied with fused multiply-adds, the small part first.
__device__ __forceinline__ void surplus_madd(double& coe, const SurplusArgs& a, int w, double x) { coe += a.wt[w] * x; }
__device__ __forceinline__ void surplus_madd(float& coe, const SurplusArgs& a, int w, float x) {
  coe = fmaf(a.wl[w], x, coe);
  coe = fmaf(a.wh[w], x, coe);
}

template <class T>
__device__ __forceinline__ void surplus_point(const SurplusArgs& a, int64_t gid) {
  const int64_t per = a.E * a.P;
  if (gid >= per * a.ncomp) return;
  const int64_t c = gid / per;
  const int64_t rem = gid - c * per;
  const int64_t e = rem / a.P;
  const int q = static_cast<int>(rem - e * a.P);
  const T* in = surplus_in<T>(a) + c * a.comp_stride;
  T coe = in[e * a.P + q];
  const int s = a.slot[e];
  if (s != 0) {
    const int pd = (q / a.stride_d) % a.p1;
    for (int ic = 0; ic < a.p1; ++ic) {
      const int anc = a.fmap[a.fbase[e] + a.anc_slot[s * a.p1 + ic]];
      if (anc < 0) { atomicOr(a.err, ERR_ANCESTOR); continue; }
      const int mm = q + (a.anc_point[s * a.p1 + ic] - pd) * a.stride_d;
      surplus_madd(coe, a, (s * a.p1 + pd) * a.p1 + ic, in[static_cast<int64_t>(anc) * a.P + mm]);
    }
  }
  if constexpr (sizeof(T) == 8) a.out[c * a.comp_stride + e * a.P + q] = coe;
  else a.out32[c * a.comp_stride + e * a.P + q] = coe;
}

}  // namespace adg
