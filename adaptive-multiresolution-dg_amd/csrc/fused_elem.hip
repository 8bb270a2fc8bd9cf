// Per-element gather kernels of the fused RK stage (thread mappings 0..3).
//
// Every thread evaluates whole terms for its outputs by walking the related
// cells of the target's fiber through the fiber map and gathering their
// values and table blocks from global memory.  Terms are summed in ascending
// source-slot order with no FMA (FP64), the order of the CPU oracle, so the
// results are bit-identical to it; the fiber-tile kernels of fused_tile.hip
// compute the same values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "kern.hpp"

namespace adg {
namespace {

constexpr int kBlock = 256;
// Final fused kernels: a block is (terms x outputs); each wavefront evaluates
// one term for consecutive outputs (same element, same fiber walk), the
// terms of an output are then added in order from LDS.
constexpr int kFinalBlock = 1024;
// Per-element final kernels hold a block of accumulators per thread: at most
// 512 threads per block leaves them 256 VGPRs.
constexpr int kElemBlock = 512;

inline unsigned nblocks(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// Value of one sweep at output (e, idx): coef * sum over related sources, in
// ascending slot order, of sum_k in[src][idx with dd <- k] * M[s*PF1+k][st*pt1+idx[dd]].
// Loads of up to four sources are issued before their terms are added in order.
template <int PF1, class T = double>
__device__ __forceinline__ T term_value(const TermDev& t, int64_t e, const int* idx, int D, int cap) {
  using Ops = SweepOps<T>;
  int base = 0;
  for (int d = 0; d < D; ++d)
    if (d != t.dd) base += idx[d] * t.sin[d];
  const int sdd = t.sin[t.dd];
  const int st = t.slot[e];
  const int fb = t.fbase[e];
  const int64_t Mc = static_cast<int64_t>(st) * t.ns * PF1 * t.pt1 + idx[t.dd];
  int maxl = t.fmaxl[e];
  if (cap >= 0) {  // coarse grid, as in k_sweep
    const int ls = t.lsum[e];
    if (ls > cap) return T(0);
    maxl = min(maxl, cap - ls + slot_level(st));
  }
  T acc = T(0);
  // sources are gathered NB at a time so that their loads are in flight together
  constexpr int NB = (PF1 <= 4) ? 4 : 2;
  int pend[NB];
  int np = 0;
  auto flush = [&]() {
    int src[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) src[j] = (j < np) ? t.fmap[fb + pend[j]] : -1;
    T xv[NB][PF1];
    typename Ops::M mv[NB][PF1];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int64_t x = static_cast<int64_t>(src[j] < 0 ? 0 : src[j]) * t.Tin + base;
      const int64_t Mr = Mc + (j < np ? pend[j] : 0) * PF1 * t.pt1;
#pragma unroll
      for (int kk = 0; kk < PF1; ++kk) {
        xv[j][kk] = (src[j] >= 0) ? Ops::in(t, x + kk * sdd) : T(0);
        mv[j][kk] = (src[j] >= 0) ? Ops::mat(t, Mr + kk * t.pt1) : Ops::zero();
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (src[j] >= 0) {
#pragma unroll
        for (int kk = 0; kk < PF1; ++kk) Ops::madd(acc, xv[j][kk], mv[j][kk]);
      }
    np = 0;
  };
  for_each_present(st, maxl, t.flx != 0, t.lu, t.fbits, fb, [&](int s) {
    pend[np++] = s;
    if (np == NB) flush();
  });
  if (np > 0) flush();
  return acc * T(term_coef(t, e));
}

__device__ __forceinline__ void unflatten(int o, const int* sz, int D, int* idx) {
  for (int d = D - 1; d >= 0; --d) {
    idx[d] = o % sz[d];
    o /= sz[d];
  }
}

// Per-output form of k_phase: thread per (task, element, output).
template <int PF1, class T>
__global__ __launch_bounds__(kBlock) void k_phase_pt(const TaskDev* tasks, int ntask, int64_t total, int D, int cap) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  int j = 0;
  while (j + 1 < ntask && gid >= tasks[j + 1].begin_v[0]) ++j;
  const TaskDev& tk = tasks[j];
  const int64_t local = gid - tk.begin_v[0];
  const int64_t e = local / tk.Tout;
  const int o = static_cast<int>(local - e * tk.Tout);
  int idx[kMaxDimK];
  unflatten(o, tk.osz, D, idx);
  SweepOps<T>::out(tk)[local] = term_value<PF1, T>(tk.t, e, idx, D, cap);
}

// Per-output form of k_up_final: threads are (point, term) pairs, one term per
// wavefront, and the point's first thread adds the terms in order.
template <int PF1, class T>
__global__ __launch_bounds__(kFinalBlock) void k_up_final_pt(const TermDev* terms, int nterm, int per_block, int64_t E,
                                                           int P, int p1, int D, T* up, T* fp,
                                                           const double* vcoef, int kind, double s, int ncomp, int cap) {
  __shared__ double vals[kFinalBlock];
  const int t = threadIdx.x / per_block, lo = threadIdx.x - t * per_block;
  const int64_t gid = static_cast<int64_t>(xcd_block()) * per_block + lo;
  const bool live = gid < E * P;
  int64_t e = 0;
  int q = 0;
  if (live) {
    e = gid / P;
    q = static_cast<int>(gid - e * P);
  }
  double v = 0.;
  if (live && t < nterm) {
    int sz[kMaxDimK], idx[kMaxDimK];
    for (int d = 0; d < D; ++d) sz[d] = p1;
    unflatten(q, sz, D, idx);
    v = term_value<PF1, T>(terms[t], e, idx, D, cap);
  }
  vals[threadIdx.x] = v;
  __syncthreads();
  if (!live || t != 0) return;
  double u = 0.;
  for (int j = 0; j < nterm; ++j) u += vals[j * per_block + lo];
  const T uu = T(u);
  up[gid] = uu;
  flux_point<T>(kind, uu, s, vcoef, fp, gid, E * P, ncomp);
}

// Per-output form of k_rhs_final: threads are (DoF, term) pairs.
template <int PFM, int PFK, class T>
__global__ __launch_bounds__(kFinalBlock) void k_rhs_final_pt(const TermDev* terms, int nterm, const TermDev* pens,
                                                            int npen, int per_block, int64_t E, int K, int k1, int D,
                                                            double* rhs, const double* u, const double* tn,
                                                            double* u_next, int scheme, int stage, double dt,
                                                            const double* dtv, const int32_t* inst, int cap, int32_t* err) {
  __shared__ double vals[kFinalBlock];
  const int t = threadIdx.x / per_block, lo = threadIdx.x - t * per_block;
  const int64_t gid = static_cast<int64_t>(xcd_block()) * per_block + lo;
  const bool live = gid < E * K;
  int64_t e = 0;
  int q = 0;
  if (live) {
    e = gid / K;
    q = static_cast<int>(gid - e * K);
  }
  double v = 0.;
  if (live && t < nterm + npen) {
    int sz[kMaxDimK], idx[kMaxDimK];
    for (int d = 0; d < D; ++d) sz[d] = k1;
    unflatten(q, sz, D, idx);
    v = (t < nterm) ? double(term_value<PFM, T>(terms[t], e, idx, D, cap))
                    : term_value<PFK, double>(pens[t - nterm], e, idx, D, cap);
  }
  vals[threadIdx.x] = v;
  __syncthreads();
  if (!live || t != 0) return;
  double r = 0.;
  for (int j = 0; j < nterm + npen; ++j) r += vals[j * per_block + lo];
  rhs[gid] = r;
  if (dtv) dt = dtv[inst[e]];
  const double un = rk_combine(scheme, stage, dt, tn[gid], u[gid], r);
  u_next[gid] = un;
  if (!isfinite(un)) atomicOr(err, ERR_NONFINITE);
}


// A thread evaluates one term for up to kRC "rest" indices of one element
// (the indices of the dimensions other than the swept one) and every output
// index of the swept dimension: one walk over the fiber serves all of them.
constexpr int kRC = 4;

// Input and output offsets of rest index r (row-major over the dimensions
// other than dd) in tensors with input sizes isz / strides si and output
// strides so.
__device__ __forceinline__ void rest_offsets(int r, int dd, int D, const int* isz, const int* si, const int* so,
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

__device__ __forceinline__ void strides_dev(const int* sz, int D, int* st) {
  st[D - 1] = 1;
  for (int d = D - 2; d >= 0; --d) st[d] = st[d + 1] * sz[d + 1];
}

// acc[r][i] = sum over related sources of element e, in ascending slot order,
// of sum_k in[src][bin[r] + k*sdd] * M(s*PF1+k, st*PF1+i), for r < nr <= RC.
// RC (1 or kRC) sizes the register arrays.  Returns false when e lies above
// the coarse-grid cap (its outputs are zero).
template <int PF1, int PT1, class T, int RC>
__device__ __forceinline__ bool term_block(const TermDev& t, int64_t e, const int* bin, int nr, int cap,
                                           T (&acc)[RC][PT1]) {
  using Ops = SweepOps<T>;
#pragma unroll
  for (int r = 0; r < RC; ++r)
#pragma unroll
    for (int i = 0; i < PT1; ++i) acc[r][i] = T(0);
  const int st = t.slot[e];
  const int fb = t.fbase[e];
  int maxl = t.fmaxl[e];
  if (cap >= 0) {  // coarse grid, as in k_sweep
    const int ls = t.lsum[e];
    if (ls > cap) return false;
    maxl = min(maxl, cap - ls + slot_level(st));
  }
  const int sdd = t.sin[t.dd];
  const int64_t Mst = static_cast<int64_t>(st) * t.ns * PF1 * PT1;
  for_each_present(st, maxl, t.flx != 0, t.lu, t.fbits, fb, [&](int s) {
    const int src = t.fmap[fb + s];
    if (src < 0) return;
    const int64_t Mr = Mst + s * PF1 * PT1;
    typename Ops::M m[PF1][PT1];
#pragma unroll
    for (int k = 0; k < PF1; ++k)
#pragma unroll
      for (int i = 0; i < PT1; ++i) m[k][i] = Ops::mat(t, Mr + k * PT1 + i);
    const int64_t xs = static_cast<int64_t>(src) * t.Tin;
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      if (r >= nr) break;
      T xv[PF1];
#pragma unroll
      for (int k = 0; k < PF1; ++k) xv[k] = Ops::in(t, xs + bin[r] + k * sdd);
#pragma unroll
      for (int i = 0; i < PT1; ++i)
#pragma unroll
        for (int k = 0; k < PF1; ++k) Ops::madd(acc[r][i], xv[k], m[k][i]);
    }
  });
  return true;
}

// Evaluates term t for (element e, chunk c) and stores coef * value at
// out[bout + i*so_dd] (zero above the coarse-grid cap).
template <int PF1, int PT1, class T, class O, int RC>
__device__ __forceinline__ void term_chunk_store(const TermDev& t, int64_t e, int c, int rc, int D, const int* so,
                                                 int cap, O* out) {
  const int R = t.Tin / t.isz[t.dd];
  const int r0 = c * rc;
  const int nr = min(rc, R - r0);
  int bin[RC], bout[RC];
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    bin[r] = 0;
    bout[r] = 0;
    if (r < nr) rest_offsets(r0 + r, t.dd, D, t.isz, t.sin, so, bin[r], bout[r]);
  }
  T acc[RC][PT1];
  const bool live = term_block<PF1, PT1, T, RC>(t, e, bin, nr, cap, acc);
  const int sod = so[t.dd];
  const double cf = term_coef(t, e);
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    if (r >= nr) break;
#pragma unroll
    for (int i = 0; i < PT1; ++i) out[bout[r] + i * sod] = live ? O(acc[r][i] * T(cf)) : O(0);
  }
}

// First sweeps of the partial sums of one phase: thread per (task, element,
// chunk of rc rest indices); variant v (1: rc = 1, 2: rc = kRC) selects the
// work-item offsets.
template <int PF1, int PT1, class T, int RC>
__global__ __launch_bounds__(kBlock) void k_phase(const TaskDev* tasks, int ntask, int64_t total, int D, int cap,
                                                  int v, int rc) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  int j = 0;
  while (j + 1 < ntask && gid >= tasks[j + 1].begin_v[v]) ++j;
  const TaskDev& tk = tasks[j];
  const int nch = (tk.R + rc - 1) / rc;
  const int64_t local = gid - tk.begin_v[v];
  const int64_t e = local / nch;
  const int c = static_cast<int>(local - e * nch);
  int so[kMaxDimK];
  strides_dev(tk.osz, D, so);
  term_chunk_store<PF1, PT1, T, T, RC>(tk.t, e, c, rc, D, so, cap, SweepOps<T>::out(tk) + e * tk.Tout);
}

// Last sweep of every eval_up partial sum, accumulated in ordering order into
// the point value, followed by the point flux (LagrInterpolation::eval_fp_Lag).
// A block holds EB elements; threads are (term, element, chunk), term-major,
// and write their values to LDS; the point values then add the terms in order.
template <int PF1, int PT1, class T, int RC>
__global__ __launch_bounds__(kElemBlock) void k_up_final(const TermDev* terms, int nterm, int EB, int nch, int rc,
                                                           int64_t E, int P, int D, T* up, T* fp,
                                                           const double* vcoef, int kind, double s, int ncomp, int cap) {
  extern __shared__ double vals[];  // [nterm][EB][P]
  const int tid = threadIdx.x;
  const int per_term = EB * nch, nslot = blockDim.x / per_term;
  const int slot = tid / per_term, rem = tid - slot * per_term;
  const int el = rem / nch, c = rem - (rem / nch) * nch;
  const int64_t e0 = static_cast<int64_t>(xcd_block()) * EB;
  int so[kMaxDimK], osz[kMaxDimK];
  for (int d = 0; d < D; ++d) osz[d] = PT1;
  strides_dev(osz, D, so);
  if (e0 + el < E)
    for (int t = slot; t < nterm; t += nslot)
      term_chunk_store<PF1, PT1, T, double, RC>(terms[t], e0 + el, c, rc, D, so, cap,
                                            vals + (static_cast<int64_t>(t) * EB + el) * P);
  __syncthreads();
  const int64_t n = E * P;
  for (int o = tid; o < EB * P; o += blockDim.x) {
    const int ol = o / P, q = o - ol * P;
    if (e0 + ol >= E) break;
    double u = 0.;
    for (int j = 0; j < nterm; ++j) u += vals[(static_cast<int64_t>(j) * EB + ol) * P + q];
    const int64_t gid = (e0 + ol) * P + q;
    const T uu = T(u);
    up[gid] = uu;
    flux_point<T>(kind, uu, s, vcoef, fp, gid, n, ncomp);
  }
}

// Last sweep of every right-hand-side partial sum and the penalty sweeps,
// accumulated in the reference order, then the RK stage update written to
// u_next (u itself is still read by other threads' penalty sweeps).  Same
// block structure as k_up_final.
template <int PFM, int PFK, class T, int RC>
__global__ __launch_bounds__(kElemBlock) void k_rhs_final(const TermDev* terms, int nterm, const TermDev* pens,
                                                            int npen, int EB, int nch, int rc, int64_t E, int K, int D,
                                                            double* rhs, const double* u, const double* tn,
                                                            double* u_next, int scheme, int stage, double dt,
                                                            const double* dtv, const int32_t* inst, int cap, int32_t* err) {
  extern __shared__ double vals[];  // [nterm + npen][EB][K]
  const int tid = threadIdx.x;
  const int per_term = EB * nch, nslot = blockDim.x / per_term;
  const int slot = tid / per_term, rem = tid - slot * per_term;
  const int el = rem / nch, c = rem - (rem / nch) * nch;
  const int64_t e0 = static_cast<int64_t>(xcd_block()) * EB;
  int so[kMaxDimK], osz[kMaxDimK];
  for (int d = 0; d < D; ++d) osz[d] = PFK;
  strides_dev(osz, D, so);
  if (e0 + el < E)
    for (int t = slot; t < nterm + npen; t += nslot) {
      double* out = vals + (static_cast<int64_t>(t) * EB + el) * K;
      if (t < nterm) term_chunk_store<PFM, PFK, T, double, RC>(terms[t], e0 + el, c, rc, D, so, cap, out);
      else term_chunk_store<PFK, PFK, double, double, RC>(pens[t - nterm], e0 + el, c, rc, D, so, cap, out);
    }
  __syncthreads();
  for (int o = tid; o < EB * K; o += blockDim.x) {
    const int ol = o / K, q = o - ol * K;
    if (e0 + ol >= E) break;
    double r = 0.;
    for (int j = 0; j < nterm + npen; ++j) r += vals[(static_cast<int64_t>(j) * EB + ol) * K + q];
    const int64_t gid = (e0 + ol) * K + q;
    rhs[gid] = r;
    const double dte = dtv ? dtv[inst[e0 + ol]] : dt;
    const double un = rk_combine(scheme, stage, dte, tn[gid], u[gid], r);
    u_next[gid] = un;
    if (!isfinite(un)) atomicOr(err, ERR_NONFINITE);
  }
}

// Outputs per block of a per-output final kernel with nterm terms: a power
// of two, at most one wavefront, with nterm * outputs <= kFinalBlock.
inline int final_outputs_per_block(int nterm) {
  int o = 64;
  while (o > 1 && o * nterm > kFinalBlock) o /= 2;
  return o;
}

// Elements per block of a per-element final kernel with nt terms, nch chunks
// per element and `outs` outputs per element (at most kElemBlock threads and
// 64 KB of LDS), and the number of term slots: threads are (slot, element,
// chunk) and a slot takes the terms slot, slot + nslot, ...
inline int final_elements_per_block(int nt, int nch, int outs) {
  if (nch > kElemBlock || nt * outs * 8 > 65536) throw std::runtime_error("final kernel: block too large");
  int eb = std::max(1, std::min(kElemBlock / (nt * nch), 65536 / (nt * outs * 8)));
  if (nch == 1 && eb >= 64) eb -= eb % 64;  // whole wavefronts per term
  return eb;
}

inline int final_term_slots(int nt, int nch, int eb) { return std::max(1, std::min(nt, kElemBlock / (eb * nch))); }
template <int PF1, int PT1, class T>
void launch_phase_t(const TaskDev* tasks, int ntask, int64_t total, int D, int cap, int v, hipStream_t s) {
  if (v == 0)
    hipLaunchKernelGGL((k_phase_pt<PF1, T>), dim3(nblocks(total)), dim3(kBlock), 0, s, tasks, ntask, total, D, cap);
  else if (v == 1)
    hipLaunchKernelGGL((k_phase<PF1, PT1, T, 1>), dim3(nblocks(total)), dim3(kBlock), 0, s, tasks, ntask, total, D,
                       cap, v, 1);
  else if (v == 3)
    hipLaunchKernelGGL((k_phase<PF1, PT1, T, 2>), dim3(nblocks(total)), dim3(kBlock), 0, s, tasks, ntask, total, D,
                       cap, v, 2);
  else
    hipLaunchKernelGGL((k_phase<PF1, PT1, T, kRC>), dim3(nblocks(total)), dim3(kBlock), 0, s, tasks, ntask, total, D,
                       cap, v, kRC);
}

template <int K1, int M1, class T>
void up_final_t(int v, const TermDev* terms, int nterm, int64_t E, int P, int D, T* up, T* fp, const double* vcoef,
                int kind, double fs, int ncomp, int cap, hipStream_t s) {
  if (v == 0) {
    const int64_t n = E * P;
    const int per_block = final_outputs_per_block(nterm);
    const unsigned nb = static_cast<unsigned>((n + per_block - 1) / per_block);
    const unsigned bs = static_cast<unsigned>(per_block * nterm);
    hipLaunchKernelGGL((k_up_final_pt<K1, T>), dim3(nb), dim3(bs), 0, s, terms, nterm, per_block, E, P, M1, D, up, fp,
                       vcoef, kind, fs, ncomp, cap);
    return;
  }
  const int rc = v == 1 ? 1 : (v == 3 ? 2 : kRC);
  const int R = P / M1;
  const int nch = (R + rc - 1) / rc;
  const int EB = final_elements_per_block(nterm, nch, P);
  const unsigned nb = static_cast<unsigned>((E + EB - 1) / EB);
  const unsigned bs = static_cast<unsigned>(final_term_slots(nterm, nch, EB) * EB * nch);
  const size_t lds = sizeof(double) * nterm * EB * P;
  if (rc == 1)
    hipLaunchKernelGGL((k_up_final<K1, M1, T, 1>), dim3(nb), dim3(bs), lds, s, terms, nterm, EB, nch, rc, E, P, D, up,
                       fp, vcoef, kind, fs, ncomp, cap);
  else if (rc == 2)
    hipLaunchKernelGGL((k_up_final<K1, M1, T, 2>), dim3(nb), dim3(bs), lds, s, terms, nterm, EB, nch, rc, E, P, D, up,
                       fp, vcoef, kind, fs, ncomp, cap);
  else
    hipLaunchKernelGGL((k_up_final<K1, M1, T, kRC>), dim3(nb), dim3(bs), lds, s, terms, nterm, EB, nch, rc, E, P, D,
                       up, fp, vcoef, kind, fs, ncomp, cap);
}

template <int M1, int K1, class T>
void rhs_final_t(int v, const TermDev* terms, int nterm, const TermDev* pens, int npen, int64_t E, int K, int D,
                 double* rhs, const double* u, const double* tn, double* u_next, int scheme, int stage, double dt,
                 const double* dtv, const int32_t* inst, int cap, int32_t* err, hipStream_t s) {
  const int nt = nterm + npen;
  if (v == 0) {
    const int64_t n = E * K;
    const int per_block = final_outputs_per_block(nt);
    const unsigned nb = static_cast<unsigned>((n + per_block - 1) / per_block);
    const unsigned bs = static_cast<unsigned>(per_block * nt);
    hipLaunchKernelGGL((k_rhs_final_pt<M1, K1, T>), dim3(nb), dim3(bs), 0, s, terms, nterm, pens, npen, per_block, E,
                       K, K1, D, rhs, u, tn, u_next, scheme, stage, dt, dtv, inst, cap, err);
    return;
  }
  const int rc = v == 1 ? 1 : (v == 3 ? 2 : kRC);
  const int R = K / K1;
  const int nch = (R + rc - 1) / rc;
  const int EB = final_elements_per_block(nt, nch, K);
  const unsigned nb = static_cast<unsigned>((E + EB - 1) / EB);
  const unsigned bs = static_cast<unsigned>(final_term_slots(nt, nch, EB) * EB * nch);
  const size_t lds = sizeof(double) * nt * EB * K;
  if (rc == 1)
    hipLaunchKernelGGL((k_rhs_final<M1, K1, T, 1>), dim3(nb), dim3(bs), lds, s, terms, nterm, pens, npen, EB, nch, rc,
                       E, K, D, rhs, u, tn, u_next, scheme, stage, dt, dtv, inst, cap, err);
  else if (rc == 2)
    hipLaunchKernelGGL((k_rhs_final<M1, K1, T, 2>), dim3(nb), dim3(bs), lds, s, terms, nterm, pens, npen, EB, nch, rc,
                       E, K, D, rhs, u, tn, u_next, scheme, stage, dt, dtv, inst, cap, err);
  else
    hipLaunchKernelGGL((k_rhs_final<M1, K1, T, kRC>), dim3(nb), dim3(bs), lds, s, terms, nterm, pens, npen, EB, nch,
                       rc, E, K, D, rhs, u, tn, u_next, scheme, stage, dt, dtv, inst, cap, err);
}

}  // namespace

template <class T>
void elem_phase(int pf1, int pt1, const TaskDev* tasks, int ntask, int64_t total, int D, int cap, int v,
                hipStream_t s) {
#define ADG_PH(F, P) if (pf1 == F && pt1 == P) { launch_phase_t<F, P, T>(tasks, ntask, total, D, cap, v, s); return; }
  ADG_PH(2, 2) ADG_PH(2, 3) ADG_PH(2, 4) ADG_PH(2, 5) ADG_PH(3, 2) ADG_PH(3, 3) ADG_PH(3, 4) ADG_PH(3, 5)
  ADG_PH(4, 2) ADG_PH(4, 3) ADG_PH(4, 4) ADG_PH(4, 5) ADG_PH(5, 2) ADG_PH(5, 3) ADG_PH(5, 4)
#undef ADG_PH
  throw std::runtime_error("phase: unsupported basis sizes");
}

template <class T>
void elem_up_final(int v, int k1, int m1, const TermDev* terms, int nterm, int64_t E, int P, int D, T* up, T* fp,
                   const double* vcoef, int kind, double fs, int ncomp, int cap, hipStream_t s) {
#define ADG_UF(K1, M1)                                                                              \
  if (k1 == K1 && m1 == M1) {                                                                       \
    up_final_t<K1, M1, T>(v, terms, nterm, E, P, D, up, fp, vcoef, kind, fs, ncomp, cap, s);        \
    return;                                                                                         \
  }
  ADG_UF(2, 2) ADG_UF(2, 3) ADG_UF(2, 4) ADG_UF(2, 5) ADG_UF(3, 2) ADG_UF(3, 3) ADG_UF(3, 4) ADG_UF(3, 5)
  ADG_UF(4, 2) ADG_UF(4, 3) ADG_UF(4, 4) ADG_UF(4, 5)
#undef ADG_UF
  throw std::runtime_error("up final: unsupported basis sizes");
}

template <class T>
void elem_rhs_final(int v, int m1, int k1, const TermDev* terms, int nterm, const TermDev* pens, int npen, int64_t E,
                    int K, int D, double* rhs, const double* u, const double* tn, double* u_next, int scheme, int stage,
                    double dt, const double* dtv, const int32_t* inst, int cap, int32_t* err, hipStream_t s) {
#define ADG_RF(M1, K1)                                                                                             \
  if (m1 == M1 && k1 == K1) {                                                                                      \
    rhs_final_t<M1, K1, T>(v, terms, nterm, pens, npen, E, K, D, rhs, u, tn, u_next, scheme, stage, dt, dtv, inst, \
                           cap, err, s);                                                                           \
    return;                                                                                                        \
  }
  ADG_RF(2, 2) ADG_RF(2, 3) ADG_RF(2, 4) ADG_RF(3, 2) ADG_RF(3, 3) ADG_RF(3, 4) ADG_RF(4, 2) ADG_RF(4, 3)
  ADG_RF(4, 4) ADG_RF(5, 2) ADG_RF(5, 3) ADG_RF(5, 4)
#undef ADG_RF
  throw std::runtime_error("rhs final: unsupported basis sizes");
}

template void elem_phase<double>(int, int, const TaskDev*, int, int64_t, int, int, int, hipStream_t);
template void elem_phase<float>(int, int, const TaskDev*, int, int64_t, int, int, int, hipStream_t);
template void elem_up_final<double>(int, int, int, const TermDev*, int, int64_t, int, int, double*, double*,
                                    const double*, int, double, int, int, hipStream_t);
template void elem_up_final<float>(int, int, int, const TermDev*, int, int64_t, int, int, float*, float*,
                                   const double*, int, double, int, int, hipStream_t);
template void elem_rhs_final<double>(int, int, int, const TermDev*, int, const TermDev*, int, int64_t, int, int,
                                     double*, const double*, const double*, double*, int, int, double, const double*,
                                     const int32_t*, int, int32_t*, hipStream_t);
template void elem_rhs_final<float>(int, int, int, const TermDev*, int, const TermDev*, int, int64_t, int, int,
                                    double*, const double*, const double*, double*, int, int, double, const double*,
                                    const int32_t*, int, int32_t*, hipStream_t);

}  // namespace adg
