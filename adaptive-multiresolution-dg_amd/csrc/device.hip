// HIP kernels of the adaptive sparse-grid DG hot path for gfx950, and the
// device context that sequences them.
//
// FP64 throughout, compiled with -ffp-contract=off so that every product and
// sum rounds exactly as the reference's x86-64 build does (no FMA).  Each
// output value accumulates its terms in ascending source-slot order, the
// order the oracle uses, so results are bit-identical to it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <numeric>
#include <thread>

#include "context.hpp"
#include "fiber.hpp"
#include "kern.hpp"
#include "surplus_dev.hpp"
#include "tile.hpp"
#include "coop.hpp"

namespace adg {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Runs fn(i) for i in [0, n) on a persistent pool of up to 15 host threads
// plus the caller (instances are independent).  Every calling host thread has
// its own pool, started on its first use, so contexts driven from different
// threads do not wait for each other's loops.
namespace {
class HostPool {
 public:
  static HostPool& get() {
    thread_local HostPool pool;
    return pool;
  }
  void run(int n, const std::function<void(int)>& fn) {
    std::lock_guard<std::mutex> one(run_mu_);
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      active_ = static_cast<int>(workers_.size());
      err_ = nullptr;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> g(mu_);
    done_cv_.wait(g, [&] { return active_ == 0; });
    fn_ = nullptr;
    if (err_) std::rethrow_exception(err_);
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }

 private:
  HostPool() {
    const int nt = std::max(0, std::min(15, static_cast<int>(std::thread::hardware_concurrency()) - 1));
    for (int t = 0; t < nt; ++t) workers_.emplace_back([this] { loop(); });
  }
  void work() {
    for (int i = next_++; i < n_; i = next_++) {
      try {
        (*fn_)(i);
      } catch (...) {
        std::lock_guard<std::mutex> g(mu_);
        if (!err_) err_ = std::current_exception();
      }
    }
  }
  void loop() {
    uint64_t seen = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      work();
      std::lock_guard<std::mutex> g(mu_);
      if (--active_ == 0) done_cv_.notify_one();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_, run_mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  int n_ = 0, active_ = 0;
  std::atomic<int> next_{0};
  uint64_t gen_ = 0;
  bool stop_ = false;
  std::exception_ptr err_;
};
}  // namespace

void parallel_for(int n, const std::function<void(int)>& fn) {
  if (n <= 8) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  HostPool::get().run(n, fn);
}

namespace {

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// One 1D sweep (FastMultiplyLU::transform_1D, source/FastMultiplyLU.cpp:436-512).
// Thread per (target element, index over the dimensions other than dd).
struct SweepArgs {
  const double* in;
  double* out;    // written when non-null
  double* accum;  // accumulated (+=) when non-null
  const double* M;
  const int32_t* slot;
  const int32_t* fbase;
  const int32_t* fmaxl;
  const int32_t* fmap;
  const uint32_t* fbits;  // occupancy bits of the fiber map
  const int32_t* lsum;    // level sum per element
  int64_t E;
  int R, Tin, Tout, sdd_in, sdd_out, ns;
  int cap;  // level-sum cap (coarse grid); negative: none
  int nrest;
  int rsz[3], rsin[3], rsout[3];
  int flx, lu;
  double coef;
  const double* coef_inst;  // per-instance coefficient (replaces coef when non-null)
  const int32_t* inst;
};

template <int PF1, int PT1>
__global__ __launch_bounds__(kBlock) void k_sweep(SweepArgs a) {
  const int64_t gid = static_cast<int64_t>(xcd_block()) * blockDim.x + threadIdx.x;
  if (gid >= a.E * a.R) return;
  const int64_t e = gid / a.R;
  int r = static_cast<int>(gid - e * a.R);
  int bin = 0, bout = 0;
  for (int i = a.nrest - 1; i >= 0; --i) {
    const int q = r % a.rsz[i];
    r /= a.rsz[i];
    bin += q * a.rsin[i];
    bout += q * a.rsout[i];
  }
  const int st = a.slot[e];
  const int fb = a.fbase[e];
  const double* Mst = a.M + static_cast<int64_t>(st) * a.ns * PF1 * PT1;
  double acc[PT1];
#pragma unroll
  for (int i = 0; i < PT1; ++i) acc[i] = 0.;
  const int64_t ob = e * a.Tout + bout;
  int maxl = a.fmaxl[e];
  if (a.cap >= 0) {
    // coarse grid (FastMultiplyLU::transform_1D_coarse_grid): a target above
    // the cap outputs zero; sources in its fiber share its other levels, so
    // the cap bounds their level in dd
    const int ls = a.lsum[e];
    if (ls > a.cap) {
#pragma unroll
      for (int i = 0; i < PT1; ++i) {
        if (a.out) a.out[ob + i * a.sdd_out] = 0.;
        if (a.accum) a.accum[ob + i * a.sdd_out] += 0.;
      }
      return;
    }
    maxl = min(maxl, a.cap - ls + slot_level(st));
  }
  for_each_present(st, maxl, a.flx != 0, a.lu, a.fbits, fb, [&](int s) {
    const int src = a.fmap[fb + s];
    if (src < 0) return;
    const double* x = a.in + static_cast<int64_t>(src) * a.Tin + bin;
    const double* Mr = Mst + s * PF1 * PT1;
#pragma unroll
    for (int kk = 0; kk < PF1; ++kk) {
      const double xv = x[kk * a.sdd_in];
#pragma unroll
      for (int i = 0; i < PT1; ++i) acc[i] += xv * Mr[kk * PT1 + i];
    }
  });
  const double cf = a.coef_inst ? a.coef_inst[a.inst[e]] : a.coef;
#pragma unroll
  for (int i = 0; i < PT1; ++i) {
    const double v = acc[i] * cf;
    if (a.out) a.out[ob + i * a.sdd_out] = v;
    if (a.accum) a.accum[ob + i * a.sdd_out] += v;
  }
}

// Point flux (LagrInterpolation::eval_fp_Lag / eval_coe_u_Lag).
__global__ __launch_bounds__(kBlock) void k_flux(const double* up, double* fp, const double* vcoef, int64_t n,
                                                  int kind, double s, int ncomp) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u = up[i];
  if (kind == 0) {
    fp[i] = u * u * 0.5;
  } else if (kind == 1) {
    fp[i] = sin(u) * s;
    fp[n + i] = cos(u) * s;
  } else {
    for (int c = 0; c < ncomp; ++c) fp[c * n + i] = vcoef[c * n + i] * u;
  }
}

// f, f' and f'' of flux component c at u: Burgers u^2/2 (kind 0) or the
// sin/cos pair scaled by s (kind 1) (FluxFunction, source/Interplation.cpp:5070-5113).
template <class T>
__device__ inline T flux_der(int kind, int c, int order, T u, T s) {
  if (kind == 0) return order == 0 ? u * u * T(0.5) : (order == 1 ? u : T(1));
  if (c == 0) return order == 0 ? sin(u) * s : (order == 1 ? cos(u) * s : -sin(u) * s);
  return order == 0 ? cos(u) * s : (order == 1 ? -sin(u) * s : -cos(u) * s);
}

// Hermite P3 point flux (HermInterpolation::eval_fp_Her_1D / eval_fp_Her_2D,
// source/Interplation.cpp:2921-3030,2045-2316).  Thread per (element, degree
// of freedom).  Degree p of a dimension is (point p % 2, derivative p / 2);
// a slope of f(u) is f'(u) times the slope of u at the same point, the cross
// slope adds f''(u) u_x u_y, in the reference's order of operations.
template <class T>
__global__ __launch_bounds__(kBlock) void k_flux_herm(const T* up, T* fp, int64_t E, int D, int kind, T s,
                                                       int ncomp) {
  const int P = D == 1 ? 4 : 16;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= E * P) return;
  const int64_t e = gid / P;
  const int q = static_cast<int>(gid - e * P);
  const T* ue = up + e * P;
  const int a = D == 1 ? q : q >> 2, b = D == 1 ? 0 : (q & 3);
  const int k0 = a >> 1, k1 = b >> 1;
  const int64_t n = E * P;
  const T uq = ue[q];
  for (int c = 0; c < ncomp; ++c) {
    T v;
    if (k0 == 0 && k1 == 0) {
      v = flux_der(kind, c, 0, uq, s);
    } else if (k1 == 0) {
      const T u0 = ue[D == 1 ? a - 2 : (a - 2) * 4 + b];
      v = T(0);
      v += flux_der(kind, c, 1, u0, s) * uq;
    } else if (k0 == 0) {
      const T u0 = ue[a * 4 + b - 2];
      v = T(0);
      v += flux_der(kind, c, 1, u0, s) * uq;
    } else {
      const T u0 = ue[(a - 2) * 4 + b - 2], ux = ue[a * 4 + b - 2], uy = ue[(a - 2) * 4 + b];
      v = T(0);
      v += flux_der(kind, c, 2, u0, s) * ux * uy;
      v += flux_der(kind, c, 1, u0, s) * uq;
    }
    fp[c * n + gid] = v;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_surplus(SurplusArgs a) {
  surplus_point<T>(a, static_cast<int64_t>(xcd_block()) * blockDim.x + threadIdx.x);
}

// One RK stage update (rk_combine); a non-finite result raises ERR_NONFINITE.
__global__ __launch_bounds__(kBlock) void k_rk(double* u, const double* tn, const double* r, int64_t n, int scheme,
                                                int stage, double dt, const double* dtv, const int32_t* inst, int K,
                                                int32_t* err) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (dtv) dt = dtv[inst[i / K]];
  const double v = rk_combine(scheme, stage, dt, tn[i], u[i], r[i]);
  u[i] = v;
  if (!isfinite(v)) atomicOr(err, ERR_NONFINITE);
}

__global__ __launch_bounds__(kBlock) void k_fill(double* p, int64_t n, double v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// u *= f elementwise (the filter's per-element damping factors, expanded per DoF).
__global__ __launch_bounds__(kBlock) void k_scale(double* u, const double* f, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) u[i] *= f[i];
}

__global__ __launch_bounds__(kBlock) void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// fmap[fbase[e] + slot[e]] = e and the matching occupancy bit: every element
// owns one entry of its fiber.
__global__ __launch_bounds__(kBlock) void k_fmap_scatter(int32_t* fmap, uint32_t* fbits, const int32_t* fbase,
                                                         const int32_t* slot, int64_t E) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int b = fbase[e] + slot[e];
  fmap[b] = static_cast<int32_t>(e);
  atomicOr(&fbits[b >> 5], 1u << (b & 31));
}

__global__ __launch_bounds__(kBlock) void k_copy(double* dst, const double* src, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Sum of squares per element (VecMultiD::norm before the final root).
__global__ __launch_bounds__(kBlock) void k_sumsq(const double* u, double* out, int64_t E, int K) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double s = 0.;
  for (int i = 0; i < K; ++i) {
    const double x = u[e * K + i];
    s += x * x;
  }
  out[e] = s;
}

// Weighted Lagrange surplus sum (DGAdaptIntp::indicator_norm_intp_Lag before the root).
struct LagrNormArgs {
  const double* c;
  const double* lmass;
  const int32_t* slot[kMaxDim];
  double* out;
  int64_t E;
  int D, P, p1;
};

__global__ __launch_bounds__(kBlock) void k_lagr_norm(LagrNormArgs a) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  double norm = 0.;
  for (int q = 0; q < a.P; ++q) {
    double val = 1.0;
    int rem = q, stride = a.P;
    for (int d = 0; d < a.D; ++d) {
      stride /= a.p1;
      const int pd = rem / stride;
      rem -= pd * stride;
      val *= a.lmass[a.slot[d][e] * a.p1 + pd];
    }
    const double x = a.c[e * a.P + q];
    val *= x * x;
    norm += val;
  }
  a.out[e] = norm;
}

// Variable coefficient at every interpolation point: out[(d*E + e)*P + q] =
// scale[inst[e]*D + d] * a1[1D point index of coordinate (d+1) mod D].
struct VcoefArgs {
  double* out;
  const double* a1;
  const int32_t* slot[kMaxDim];
  const int32_t* inst;
  const double* scale;
  int64_t E;
  int D, P, p1;
};

__global__ __launch_bounds__(kBlock) void k_vcoef(VcoefArgs a) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= a.E * a.P) return;
  const int64_t e = gid / a.P;
  const int q = static_cast<int>(gid - e * a.P);
  int idx[kMaxDim];
  int rem = q;
  for (int d = a.D - 1; d >= 0; --d) {
    idx[d] = rem % a.p1;
    rem /= a.p1;
  }
  for (int d = 0; d < a.D; ++d) {
    const int dn = (d + 1) % a.D;
    const double v = a.a1[a.slot[dn][e] * a.p1 + idx[dn]];
    a.out[(d * a.E + e) * a.P + q] = a.scale ? a.scale[a.inst[e] * a.D + d] * v : v;
  }
}

// ---------------------------------------------------------------------------
// Point values of the DG solution on a tensor grid of coordinates
// (DGSolution::val, source/DGSolution.cpp:613-625 with Element::val,
// source/Element.cpp:123-147).  For every 1D coordinate the host lists the
// cells whose closed support contains it, level by level, with the k+1 basis
// values there; a point visits every combination of its per-dimension cells,
// finds the element through an open-addressing table keyed by the packed
// slots, and adds c * prod_d v_d over the element's coefficients.
struct EvalArgs {
  const double* u;
  const uint64_t* hkey;   // table keys (kEmptyKey: free)
  const int32_t* hval;    // element index of each key
  uint64_t hmask;
  const int32_t* ncand;   // [nx] candidates of each 1D coordinate
  const int32_t* cslot;   // [nx * cm]
  const double* cval;     // [nx * cm * k1]
  double* out;            // [nx^D]
  int64_t npts;
  int nx, cm, k1, D, K;
};

constexpr uint64_t kEmptyKey = ~uint64_t{0};

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_eval_tensor(EvalArgs a) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= a.npts) return;
  int ix[kMaxDim], nc[kMaxDim], c[kMaxDim];
  int64_t rem = gid;
  for (int d = a.D - 1; d >= 0; --d) {
    ix[d] = static_cast<int>(rem % a.nx);
    rem /= a.nx;
    nc[d] = a.ncand[ix[d]];
    c[d] = 0;
  }
  double total = 0.;
  for (int d = 0; d < a.D; ++d)
    if (nc[d] == 0) { a.out[gid] = 0.; return; }
  while (true) {
    uint64_t key = 0;
    for (int d = 0; d < a.D; ++d) key |= static_cast<uint64_t>(a.cslot[ix[d] * a.cm + c[d]]) << (16 * d);
    uint64_t h = mix64(key) & a.hmask;
    int32_t e = -1;
    while (a.hkey[h] != kEmptyKey) {
      if (a.hkey[h] == key) { e = a.hval[h]; break; }
      h = (h + 1) & a.hmask;
    }
    if (e >= 0) {
      double value = 0.;
      for (int q = 0; q < a.K; ++q) {
        double val_bas = a.u[static_cast<int64_t>(e) * a.K + q];
        int r = q, stride = a.K;
        for (int d = 0; d < a.D; ++d) {
          stride /= a.k1;
          const int p = r / stride;
          r -= p * stride;
          val_bas *= a.cval[(ix[d] * a.cm + c[d]) * a.k1 + p];
        }
        value += val_bas;
      }
      total += value;
    }
    int d = a.D - 1;
    while (d >= 0 && ++c[d] == nc[d]) { c[d] = 0; --d; }
    if (d < 0) break;
  }
  a.out[gid] = total;
}

__global__ __launch_bounds__(kBlock) void k_remap(double* dst, const double* src, const int32_t* old_of_new, int64_t E,
                                                   int T) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= E * T) return;
  const int64_t e = i / T;
  const int t = static_cast<int>(i - e * T);
  const int32_t o = old_of_new[e];
  dst[i] = (o >= 0) ? src[static_cast<int64_t>(o) * T + t] : 0.;
}


template <int PF1, int PT1>
void launch_sweep_t(const SweepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_sweep<PF1, PT1>), dim3(nblocks(a.E * a.R)), dim3(kBlock), 0, s, a);
}

void launch_sweep(const SweepArgs& a, int pf1, int pt1, hipStream_t s) {
#define ADG_SW(F, T) if (pf1 == F && pt1 == T) { launch_sweep_t<F, T>(a, s); return; }
  ADG_SW(2, 2) ADG_SW(2, 3) ADG_SW(2, 4) ADG_SW(3, 2) ADG_SW(3, 3) ADG_SW(3, 4)
  ADG_SW(4, 2) ADG_SW(4, 3) ADG_SW(4, 4)
  ADG_SW(2, 5) ADG_SW(3, 5) ADG_SW(4, 5) ADG_SW(5, 2) ADG_SW(5, 3) ADG_SW(5, 4)
  ADG_SW(2, 6) ADG_SW(3, 6) ADG_SW(4, 6) ADG_SW(6, 2) ADG_SW(6, 3) ADG_SW(6, 4)
#undef ADG_SW
  throw std::runtime_error("sweep: unsupported basis sizes (Alpert degrees 1..3, Lagrange 1..5 are provided)");
}

// Dense level-ordered table (rows ns*pf1, cols ns*pt1) -> target-major cell blocks.
std::vector<double> to_blocks(const std::vector<double>& M, int ns, int pf1, int pt1) {
  std::vector<double> b(M.size());
  const size_t cols = static_cast<size_t>(ns) * pt1;
  for (int st = 0; st < ns; ++st)
    for (int sc = 0; sc < ns; ++sc)
      for (int kk = 0; kk < pf1; ++kk)
        for (int i = 0; i < pt1; ++i)
          b[((static_cast<size_t>(st) * ns + sc) * pf1 + kk) * pt1 + i] =
              M[(static_cast<size_t>(sc) * pf1 + kk) * cols + static_cast<size_t>(st) * pt1 + i];
  return b;
}

template <class T>
void upload(DevBuf<T>& b, const std::vector<T>& h, hipStream_t s) {
  b.ensure(std::max<size_t>(h.size(), 1));
  if (!h.empty()) hip_check(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

}  // namespace

std::vector<std::vector<int>> transform_orders(int D, std::vector<std::vector<int>>& lus) {
  // FastMultiplyLU::generate_transform_order (source/FastMultiplyLU.cpp:614-664)
  std::vector<std::vector<int>> dims;
  lus.clear();
  if (D == 1) { dims.push_back({0}); lus.push_back({LU_FULL}); return dims; }
  const int n = 1 << (D - 1);
  for (int num = 0; num < n; ++num) {
    std::vector<int> v(D);
    for (int b = 0; b < D - 1; ++b) v[b] = 2 * ((num >> (D - 2 - b)) & 1);
    v[D - 1] = 1;
    std::vector<int> order(D);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return v[a] < v[b]; });
    std::vector<int> lu(D);
    for (int i = 0; i < D; ++i) {
      const int val = v[order[i]];
      lu[i] = (val == 0) ? LU_L : (val == 1 ? LU_FULL : LU_U);
    }
    dims.push_back(order);
    lus.push_back(lu);
  }
  return dims;
}

Context::Context(int dim, int k_, int m_, int nmax_)
    : D(dim), k(k_), m(m_), nmax(nmax_),
      alpt(BasisKind::Alpert, k_, nmax_), lagr(BasisKind::Lagrange, m_, nmax_) {
  if (D < 1 || D > kMaxDim) throw std::runtime_error("dimension must be 1..4");
  if (k < 1 || k > 3 || m < 1 || m > 5) throw std::runtime_error("degrees must be k in 1..3 and m in 1..5");
  if (nmax < 1 || nmax > 13) throw std::runtime_error("maximum level must be 1..13");
  K = 1; P = 1;
  for (int d = 0; d < D; ++d) { K *= k + 1; P *= m + 1; }
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  na = alpt.size();
  nl = lagr.size();
  OperatorTables aa(alpt, alpt), la(lagr, alpt);
  const int ns = 1 << nmax;
  upload(t_pt, to_blocks(alpt_at_lagr_points(alpt, lagr), ns, k + 1, m + 1), stream);
  upload(t_la_uv, to_blocks(la.u_v, ns, m + 1, k + 1), stream);
  upload(t_la_uvx, to_blocks(la.u_vx, ns, m + 1, k + 1), stream);
  std::vector<double> flx(la.ulft_vjp.size());
  for (size_t i = 0; i < flx.size(); ++i) flx[i] = la.ulft_vjp[i] + la.urgt_vjp[i];
  upload(t_la_flx, to_blocks(flx, ns, m + 1, k + 1), stream);
  upload(t_aa[0], to_blocks(aa.u_v, ns, k + 1, k + 1), stream);
  upload(t_aa[1], to_blocks(aa.u_vx, ns, k + 1, k + 1), stream);
  upload(t_aa[2], to_blocks(aa.ulft_vjp, ns, k + 1, k + 1), stream);
  upload(t_aa[3], to_blocks(aa.urgt_vjp, ns, k + 1, k + 1), stream);
  upload(t_aa[4], to_blocks(aa.ujp_vjp, ns, k + 1, k + 1), stream);
  upload(t_lmass, lagr_mass_diag(lagr), stream);
  SurplusTables sur(lagr);
  upload(t_anc_slot, sur.anc_slot, stream);
  upload(t_anc_point, sur.anc_point, stream);
  upload(t_wt, sur.wt, stream);
  err_flag.ensure(1);
  hip_check(hipMemcpyAsync(err_flag.p, std::vector<int32_t>{0}.data(), sizeof(int32_t), hipMemcpyHostToDevice, stream),
            "upload");
  sync();
}

void Context::set_interpolation(int family) {
  if (family != 0 && family != 1) throw std::runtime_error("unknown interpolation family");
  if (family == interp) return;
  if (family == 1 && !herm) {
    if (m != 3) throw std::runtime_error("Hermite interpolation needs interpolation degree m = 3");
    std::unique_ptr<Basis1D> hb(new Basis1D(BasisKind::Hermite, 3, nmax));
    OperatorTables ha(*hb, alpt);
    const int ns = 1 << nmax;
    upload(alt_pt, to_blocks(alpt_at_herm_points(alpt, *hb), ns, k + 1, m + 1), stream);
    upload(alt_uv, to_blocks(ha.u_v, ns, m + 1, k + 1), stream);
    upload(alt_uvx, to_blocks(ha.u_vx, ns, m + 1, k + 1), stream);
    std::vector<double> flx(ha.ulft_vjp.size());
    for (size_t i = 0; i < flx.size(); ++i) flx[i] = ha.ulft_vjp[i] + ha.urgt_vjp[i];
    upload(alt_flx, to_blocks(flx, ns, m + 1, k + 1), stream);
    SurplusTables hs(*hb);
    upload(alt_anc_slot, hs.anc_slot, stream);
    upload(alt_anc_point, hs.anc_point, stream);
    upload(alt_wt, hs.wt, stream);
    sync();
    herm = std::move(hb);
  }
  t_pt.swap(alt_pt);
  t_la_uv.swap(alt_uv);
  t_la_uvx.swap(alt_uvx);
  t_la_flx.swap(alt_flx);
  t_wt.swap(alt_wt);
  t_anc_slot.swap(alt_anc_slot);
  t_anc_point.swap(alt_anc_point);
  interp = family;
}

struct Context::FusedPlan {
  int64_t version = -1;
  int interp = -1;
  int precision = -1;
  const float* fucoe32 = nullptr;
  const double *u = nullptr, *tn = nullptr, *fucoe = nullptr, *tmp = nullptr, *pen_coef = nullptr;
  double* unext = nullptr;
  double* tv = nullptr;
  const int32_t *perm0 = nullptr, *pre0 = nullptr;
  int flux_kind = -1;
  double alpha[kMaxDim] = {0, 0, 0, 0};
  std::vector<int> up_phase_ntask, rhs_phase_ntask;
  std::vector<std::array<int64_t, 4>> up_phase_total, rhs_phase_total;  // work items per per-element mapping
  int n_up_terms = 0, n_rhs_terms = 0, n_pen = 0;
  // algorithmic bytes per launch (8 B x elements x (inputs + outputs))
  std::vector<double> up_phase_bytes, rhs_phase_bytes;
  double up_final_bytes = 0, rhs_final_bytes = 0;
  DevBuf<TaskDev> up_tasks, rhs_tasks;
  DevBuf<TermDev> up_terms, rhs_terms, pens;
  // fiber-tile form of the last sweeps: tasks writing one value buffer per
  // term (tile_tv: eval_up, then right-hand-side terms; tile_pv: penalties)
  DevBuf<TaskDev> up_final_tasks, rhs_final_tasks, pen_tasks;
  // rest indices of the tasks of each launch (tile LDS sizing)
  std::vector<int> up_phase_R, rhs_phase_R;
  int up_final_R = 0, rhs_final_R = 0, pen_R = 0;
};

Context::~Context() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (StepGraph& g : step_graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  for (ProfRec& r : prof_pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (hipEvent_t e : prof_pool) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void Context::sync() { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

void Context::prof_begin(int cat, double bytes) {
  if (!prof_on) return;
  ProfRec r{};
  for (hipEvent_t* ev : {&r.a, &r.b}) {
    if (prof_pool.empty()) hip_check(hipEventCreate(ev), "hipEventCreate");
    else { *ev = prof_pool.back(); prof_pool.pop_back(); }
  }
  r.cat = cat;
  r.bytes = bytes;
  hip_check(hipEventRecord(r.a, stream), "hipEventRecord");
  prof_pending.push_back(r);
}

void Context::prof_end() {
  if (!prof_on) return;
  hip_check(hipEventRecord(prof_pending.back().b, stream), "hipEventRecord");
  if (prof_pending.size() >= 4096) prof_collect();
}

void Context::prof_collect() {
  if (prof_pending.empty()) return;
  sync();
  for (ProfRec& r : prof_pending) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
    prof_ms[r.cat] += ms;
    prof_bytes[r.cat] += r.bytes;
    prof_count[r.cat] += 1;
    prof_pool.push_back(r.a);
    prof_pool.push_back(r.b);
  }
  prof_pending.clear();
}

void Context::prof_reset() {
  prof_collect();
  for (int i = 0; i < PROF_NCAT; ++i) { prof_ms[i] = 0; prof_bytes[i] = 0; prof_count[i] = 0; }
  for (double& v : host_ms) v = 0;
}

SweepTable Context::table(int id) {
  const int ns = 1 << nmax;
  switch (id) {
    case 0: return {t_pt.p, ns};
    case 1: return {t_la_uv.p, ns};
    case 2: return {t_la_uvx.p, ns};
    case 3: return {t_la_flx.p, ns};
    case 4: case 5: case 6: case 7: case 8: return {t_aa[id - 4].p, ns};
  }
  throw std::runtime_error("unknown table id");
}

void Context::host_union() {
  const int64_t n = E();
  uslot.resize(n * D);
  uhash.resize(n);
  uinst.resize(n);
  inst_lmax.assign(static_cast<size_t>(B()) * D, 0);
  parallel_for(B(), [&](int i) {
    const HostGrid& g = grids[i];
    std::copy(g.slot.begin(), g.slot.end(), uslot.begin() + off[i] * D);
    std::copy(g.hash.begin(), g.hash.end(), uhash.begin() + off[i]);
    std::fill(uinst.begin() + off[i], uinst.begin() + off[i + 1], i);
    for (int64_t e = 0; e < g.E; ++e)
      for (int d = 0; d < D; ++d)
        inst_lmax[static_cast<size_t>(i) * D + d] = std::max(inst_lmax[static_cast<size_t>(i) * D + d], g.level(e, d));
  });
}

void Context::upload_grid() {
  HostTimer timer(*this, HOST_UPLOAD);
  // union of the instance tables: element indices and fiber ids are shifted by
  // the instance offsets, so fibers stay within their instance
  const int64_t n = E();
  host_union();
  std::vector<int32_t> dsl(n), fb(n), fml(n), ls(n, 0);
  parallel_for(B(), [&](int i) {
    const HostGrid& g = grids[i];
    for (int64_t e = 0; e < g.E; ++e)
      for (int d = 0; d < D; ++d) ls[off[i] + e] += level_of_slot(g.slot[e * D + d]);
  });
  for (int d = 0; d < D; ++d) {
    std::vector<int64_t> fib0(B() + 1, 0);  // first fiber of every instance
    for (int i = 0; i < B(); ++i) fib0[i + 1] = fib0[i] + grids[i].nfib[d];
    const int64_t fib_off = fib0[B()];
    if ((fib_off << nmax) > (int64_t{1} << 31) - 1) throw std::runtime_error("fiber map exceeds int32 range");
    parallel_for(B(), [&](int i) {
      const HostGrid& g = grids[i];
      const int32_t shift = static_cast<int32_t>(fib0[i] << nmax);
      for (int64_t e = 0; e < g.E; ++e) {
        dsl[off[i] + e] = g.dslot[d][e];
        fb[off[i] + e] = g.fbase[d][e] + shift;
        fml[off[i] + e] = g.fmaxl[d][e];
      }
    });
    upload(g_slot[d], dsl, stream);
    upload(g_fbase[d], fb, stream);
    upload(g_fmaxl[d], fml, stream);
    // dense fiber map, scattered on the device: fmap[fbase[e] + slot[e]] = e
    const int64_t nmap = std::max<int64_t>(fib_off << nmax, 1);
    g_fmap[d].ensure(static_cast<size_t>(nmap));
    hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(nmap)), dim3(kBlock), 0, stream, g_fmap[d].p, nmap, -1);
    const int64_t nwords = (nmap + 31) / 32;
    // one zero word past the end: the position of the cell after the last
    // fiber is read from it
    g_fbits[d].ensure(static_cast<size_t>(nwords + 1));
    hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(nwords + 1)), dim3(kBlock), 0, stream, g_fbits[d].p, nwords + 1, 0u);
    fbits_words[d] = nwords;
    if (n > 0)
      hipLaunchKernelGGL(k_fmap_scatter, dim3(nblocks(n)), dim3(kBlock), 0, stream, g_fmap[d].p, g_fbits[d].p,
                         g_fbase[d].p, g_slot[d].p, n);
  }
  upload(g_inst, uinst, stream);
  upload(g_lsum, ls, stream);
  std::vector<uint64_t> uk(n);
  for (int64_t e = 0; e < n; ++e)
    uk[e] = (static_cast<uint64_t>(uinst[e]) << 32) | static_cast<uint32_t>(uhash[e]);
  upload(g_ukey, uk, stream);
  maxfib = 1;
  for (const HostGrid& g : grids)
    for (int dd = 0; dd < D; ++dd) maxfib = std::max(maxfib, g.maxfib[dd]);
}

void Context::ensure_fields() {
  const size_t e = static_cast<size_t>(std::max<int64_t>(E(), 1));
  const int mx = std::max(k, m) + 1;
  int T = 1;
  for (int d = 0; d < D; ++d) T *= mx;
  for (DevBuf<double>* b : {&u, &tn, &pred, &rhs}) b->ensure(e * K);
  for (DevBuf<double>* b : {&up, &uintp}) b->ensure(e * P);
  for (DevBuf<double>* b : {&fp, &fucoe, &interA, &interB, &vcoef}) b->ensure(e * P * D);
  tmpA.ensure(e * T);
  tmpB.ensure(e * T);
  sumsq.ensure(e);
}

void Context::remap_field(DevBuf<double>& f, int per_elem, const std::vector<int32_t>& old_of_new, int64_t old_E) {
  remap_field_dev(f, per_elem, static_cast<int64_t>(old_of_new.size()), old_E);
}

void Context::remap_field_dev(DevBuf<double>& f, int per_elem, int64_t newE, int64_t old_E) {
  remap_tmp.ensure(std::max<int64_t>(old_E * per_elem, 1));
  if (old_E > 0)
    hipLaunchKernelGGL(k_copy, dim3(nblocks(old_E * per_elem)), dim3(kBlock), 0, stream, remap_tmp.p, f.p,
                       old_E * per_elem);
  f.ensure(std::max<int64_t>(newE * per_elem, 1));
  if (newE > 0)
    hipLaunchKernelGGL(k_remap, dim3(nblocks(newE * per_elem)), dim3(kBlock), 0, stream, f.p, remap_tmp.p,
                       g_remap.p, newE, per_elem);
}

void Context::set_grid_slots(const std::vector<int32_t>& slots, bool keep_fields) {
  set_grids(std::vector<std::vector<int32_t>>{slots}, keep_fields);
}

void Context::set_grids(const std::vector<std::vector<int32_t>>& slots, bool keep_fields,
                        const std::vector<char>* changed) {
  ensure_host_grid();
  HostTimer timer(*this, HOST_GRID);
  const int nb = static_cast<int>(slots.size());
  // instances whose element set did not change keep their table
  auto reuse = [&](int i) { return changed && nb == B() && !(*changed)[i]; };
  std::vector<HostGrid> ng(nb);
  {
    HostTimer tb(*this, HOST_BUILD);
    parallel_for(nb, [&](int i) {
      if (!reuse(i)) ng[i].build(D, nmax, slots[i], hk);
    });
  }
  std::vector<int64_t> noff(nb + 1, 0);
  for (int i = 0; i < nb; ++i) noff[i + 1] = noff[i] + (reuse(i) ? grids[i].E : ng[i].E);
  const int64_t newE = noff[nb];
  if (keep_fields && E() > 0 && nb == B()) {
    std::vector<int32_t> m(newE, -1);
    parallel_for(nb, [&](int i) {
      if (reuse(i)) {
        for (int64_t j = 0; j < grids[i].E; ++j) m[noff[i] + j] = static_cast<int32_t>(off[i] + j);
        return;
      }
      const std::vector<int32_t> mi = map_new_to_old(grids[i], ng[i]);
      for (int64_t j = 0; j < ng[i].E; ++j) m[noff[i] + j] = mi[j] < 0 ? -1 : static_cast<int32_t>(off[i] + mi[j]);
    });
    upload(g_remap, m, stream);
    const int64_t oldE = E();
    remap_field(u, K, m, oldE);
    remap_field(pred, K, m, oldE);
    // host-side interpolation state follows the grid
    if (!up_host.empty()) {
      std::vector<double> nu(static_cast<size_t>(newE) * P, 0.);
      std::vector<char> na_(newE, 1);
      for (int64_t i = 0; i < newE; ++i)
        if (m[i] >= 0) {
          std::copy(&up_host[static_cast<size_t>(m[i]) * P], &up_host[static_cast<size_t>(m[i]) * P] + P, &nu[i * P]);
          na_[i] = 0;
        }
      up_host.swap(nu);
      new_add.swap(na_);
    }
  } else {
    up_host.clear();
    new_add.assign(newE, 1);
  }
  for (int i = 0; i < nb; ++i)
    if (reuse(i)) ng[i] = std::move(grids[i]);
  grids = std::move(ng);
  off = noff;
  ++grid_version;
  upload_grid();
  ensure_fields();
  if (!keep_fields) {
    for (DevBuf<double>* b : {&u, &tn, &pred, &rhs})
      hipLaunchKernelGGL(k_fill, dim3(nblocks(std::max<int64_t>(E() * K, 1))), dim3(kBlock), 0, stream, b->p,
                         E() * K, 0.);
  }
  // variable coefficients depend on the point set
  if (is_varcoef(flux_kind)) compute_vcoef();
}

void Context::upload_inst_dt() {
  upload(d_inst_dt, inst_dt, stream);
  use_inst_dt = true;
}

void Context::sweep(const double* in, const int* sin, double* out, double* accum, const int* sout, int dd,
                    SweepTable tab, bool flx, int lu, double coef, int pf, int pt, const double* coef_inst) {
  if (E() == 0) return;
  SweepArgs a{};
  a.in = in; a.out = out; a.accum = accum; a.M = tab.dev; a.ns = tab.ns;
  a.slot = g_slot[dd].p; a.fbase = g_fbase[dd].p; a.fmaxl = g_fmaxl[dd].p; a.fmap = g_fmap[dd].p;
  a.fbits = g_fbits[dd].p;
  a.lsum = g_lsum.p; a.cap = lsum_cap;
  a.E = E();
  int stin[kMaxDim], stout[kMaxDim];
  stin[D - 1] = 1; stout[D - 1] = 1;
  for (int d = D - 2; d >= 0; --d) { stin[d] = stin[d + 1] * sin[d + 1]; stout[d] = stout[d + 1] * sout[d + 1]; }
  a.Tin = stin[0] * sin[0];
  a.Tout = stout[0] * sout[0];
  a.sdd_in = stin[dd];
  a.sdd_out = stout[dd];
  a.nrest = 0;
  a.R = 1;
  for (int d = 0; d < D; ++d) {
    if (d == dd) continue;
    a.rsz[a.nrest] = sin[d];
    a.rsin[a.nrest] = stin[d];
    a.rsout[a.nrest] = stout[d];
    a.R *= sin[d];
    ++a.nrest;
  }
  a.flx = flx ? 1 : 0;
  a.lu = lu;
  a.coef = coef;
  a.coef_inst = coef_inst;
  a.inst = g_inst.p;
  if (sin[dd] != pf + 1 || sout[dd] != pt + 1) throw std::runtime_error("sweep: size mismatch in swept dimension");
  prof_begin(PROF_SWEEP, 8.0 * static_cast<double>(a.E) * (a.Tin + a.Tout));
  launch_sweep(a, pf + 1, pt + 1, stream);
  prof_end();
}

void Context::partial_sums(const SweepTable* tabs, const int* flx, int pf, int pt, double coef, const double* src,
                           double* dst) {
  std::vector<std::vector<int>> lus;
  const std::vector<std::vector<int>> dims = transform_orders(D, lus);
  for (size_t o = 0; o < dims.size(); ++o) {
    int sz[kMaxDim];
    for (int d = 0; d < D; ++d) sz[d] = pf + 1;
    const double* cur = src;
    for (int s = 0; s < D; ++s) {
      const int dd = dims[o][s];
      int sin[kMaxDim], sout[kMaxDim];
      for (int d = 0; d < D; ++d) sin[d] = sz[d];
      sz[dd] = pt + 1;
      for (int d = 0; d < D; ++d) sout[d] = sz[d];
      const bool last = (s == D - 1);
      double* out = last ? nullptr : ((s % 2 == 0) ? tmpA.p : tmpB.p);
      sweep(cur, sin, out, last ? dst : nullptr, sout, dd, tabs[dd], flx[dd] != 0, lus[o][s], s == 0 ? coef : 1.0,
            pf, pt);
      cur = out;
    }
  }
}

void Context::eval_up() {
  hipLaunchKernelGGL(k_fill, dim3(nblocks(E() * P)), dim3(kBlock), 0, stream, up.p, E() * P, 0.);
  SweepTable tabs[kMaxDim];
  int flx[kMaxDim];
  for (int d = 0; d < D; ++d) { tabs[d] = table(0); flx[d] = 0; }
  partial_sums(tabs, flx, k, m, 1.0, u.p, up.p);
}

void Context::compute_vcoef() {
  // a_d(x) = c_d (1 + s*sin(2 pi x_{(d+1) mod D})) at every Lagrange point.
  // The factor 1 + s*sin(..) depends on one coordinate only, so it is
  // tabulated on the host for the 2^NMAX (m+1) points of a 1D grid (with the
  // oracle's libm) and gathered on the device; c_d is the instance's scale
  // (ADG_FLUX_VARCOEF_SCALED), absent for ADG_FLUX_VARCOEF.
  not_recording("point coefficients");
  const int p1 = m + 1;
  const int npt = (1 << nmax) * p1;
  std::vector<double> a1(npt);
  for (int c = 0; c < npt; ++c) a1[c] = 1. + flux_param * std::sin(2. * PI * lagr.point(c));
  upload(vcoef_1d, a1, stream);
  vcoef_version = grid_version;
  const bool scaled = flux_kind == ADG_FLUX_VARCOEF_SCALED_ID;
  if (scaled) {
    if (static_cast<int>(inst_scale.size()) != B() * D) {
      inst_scale.assign(static_cast<size_t>(B()) * D, 1.);
    }
    upload(d_inst_scale, inst_scale, stream);
  }
  ensure_fields();
  VcoefArgs a{};
  a.out = vcoef.p; a.a1 = vcoef_1d.p; a.inst = g_inst.p; a.scale = scaled ? d_inst_scale.p : nullptr;
  for (int dd = 0; dd < D; ++dd) a.slot[dd] = g_slot[dd].p;
  a.E = E(); a.D = D; a.P = P; a.p1 = p1;
  if (E() > 0) hipLaunchKernelGGL(k_vcoef, dim3(nblocks(E() * P)), dim3(kBlock), 0, stream, a);
}

namespace {
void check_herm_flux(int kind, int D) {
  if (D > 2) throw std::runtime_error("Hermite flux interpolation is one- or two-dimensional");
  if (kind != 0 && kind != 1) throw std::runtime_error("Hermite flux interpolation needs a flux of u alone (kinds 0, 1)");
}
}  // namespace

void Context::flux(int kind, const double* params) {
  const double s = params ? params[0] : 0.;
  if (interp == 1) {
    check_herm_flux(kind, D);
    if (kind == 1 && D != 2) throw std::runtime_error("sin/cos flux is two-dimensional");
    flux_kind = kind;
    flux_param = s;
    same_flux = (kind == 0);
    const int ncomp = same_flux ? 1 : D;
    const int64_t n = E() * P;
    if (n == 0) return;
    prof_begin(PROF_FLUX, 8.0 * static_cast<double>(n) * (1 + ncomp));
    hipLaunchKernelGGL(k_flux_herm<double>, dim3(nblocks(n)), dim3(kBlock), 0, stream, up.p, fp.p, E(), D, kind, s,
                       ncomp);
    prof_end();
    return;
  }
  if (is_varcoef(kind) && (flux_kind != kind || flux_param != s || vcoef_version != grid_version)) {
    flux_kind = kind;
    flux_param = s;
    compute_vcoef();
  }
  flux_kind = kind;
  flux_param = s;
  same_flux = (kind == 0);
  if (kind == 1 && D != 2) throw std::runtime_error("sin/cos flux is two-dimensional");
  const int64_t n = E() * P;
  if (n == 0) return;
  const int nout = (kind == 0) ? 1 : D;
  prof_begin(PROF_FLUX, 8.0 * static_cast<double>(n) * (1 + nout + (is_varcoef(kind) ? D : 0)));
  hipLaunchKernelGGL(k_flux, dim3(nblocks(n)), dim3(kBlock), 0, stream, up.p, fp.p, vcoef.p, n, kind, s, D);
  prof_end();
}

void Context::surplus_transform(const double* in, double* out, double* sa, double* sb, int ncomp) {
  const int p1 = m + 1;
  const double* cur = in;
  int stride = P;
  for (int d = 0; d < D; ++d) {
    stride /= p1;
    double* dst = (d == D - 1) ? out : ((d % 2 == 0) ? sa : sb);
    SurplusArgs a{};
    a.in = cur; a.out = dst;
    a.slot = g_slot[d].p; a.fbase = g_fbase[d].p; a.fmap = g_fmap[d].p;
    a.anc_slot = t_anc_slot.p; a.anc_point = t_anc_point.p; a.wt = t_wt.p; a.err = err_flag.p;
    a.E = E(); a.comp_stride = E() * P; a.P = P; a.p1 = p1; a.stride_d = stride; a.ncomp = ncomp;
    if (E() > 0) {
      prof_begin(PROF_SURPLUS, 8.0 * static_cast<double>(E()) * P * ncomp * 2);
      hipLaunchKernelGGL(k_surplus<double>, dim3(nblocks(E() * P * ncomp)), dim3(kBlock), 0, stream, a);
      prof_end();
    }
    cur = dst;
  }
}

void Context::surplus_transform32(const float* in, float* out, float* sa, float* sb, int ncomp) {
  const int p1 = m + 1;
  const SplitTable& w = split_table(t_wt.p, static_cast<size_t>(1 << nmax) * p1 * p1);
  const float* cur = in;
  int stride = P;
  for (int d = 0; d < D; ++d) {
    stride /= p1;
    float* dst = (d == D - 1) ? out : ((d % 2 == 0) ? sa : sb);
    SurplusArgs a{};
    a.in32 = cur; a.out32 = dst; a.wh = w.hi.p; a.wl = w.lo.p;
    a.slot = g_slot[d].p; a.fbase = g_fbase[d].p; a.fmap = g_fmap[d].p;
    a.anc_slot = t_anc_slot.p; a.anc_point = t_anc_point.p; a.err = err_flag.p;
    a.E = E(); a.comp_stride = E() * P; a.P = P; a.p1 = p1; a.stride_d = stride; a.ncomp = ncomp;
    if (E() > 0) {
      prof_begin(PROF_SURPLUS, 4.0 * static_cast<double>(E()) * P * ncomp * 2);
      hipLaunchKernelGGL(k_surplus<float>, dim3(nblocks(E() * P * ncomp)), dim3(kBlock), 0, stream, a);
      prof_end();
    }
    cur = dst;
  }
}

// hi = fl32(M), lo = fl32(M - hi)
__global__ __launch_bounds__(kBlock) void k_split(const double* M, float* hi, float* lo, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float h = static_cast<float>(M[i]);
  hi[i] = h;
  lo[i] = static_cast<float>(M[i] - static_cast<double>(h));
}

const Context::SplitTable& Context::split_table(const double* dev, size_t n) {
  auto it = split_tabs.find(dev);
  if (it != split_tabs.end()) return *it->second;
  not_recording("split table");
  std::unique_ptr<SplitTable> t(new SplitTable());
  t->hi.ensure(n);
  t->lo.ensure(n);
  hipLaunchKernelGGL(k_split, dim3(nblocks(static_cast<int64_t>(n))), dim3(kBlock), 0, stream, dev, t->hi.p, t->lo.p,
                     static_cast<int64_t>(n));
  const SplitTable& r = *t;
  split_tabs[dev] = std::move(t);
  return r;
}

void Context::fp_to_coe() {
  const int ncomp = same_flux ? 1 : D;
  surplus_transform(fp.p, fucoe.p, interA.p, interB.p, ncomp);
}

void Context::rhs_zero() {
  if (E() > 0) hipLaunchKernelGGL(k_fill, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, rhs.p, E() * K, 0.);
}

void Context::rhs_lagr(bool flux_term) {
  for (int d = 0; d < D; ++d) {
    SweepTable tabs[kMaxDim];
    int flx[kMaxDim];
    for (int dd = 0; dd < D; ++dd) {
      if (dd == d) { tabs[dd] = flux_term ? table(3) : table(2); flx[dd] = flux_term ? 1 : 0; }
      else { tabs[dd] = table(1); flx[dd] = 0; }
    }
    const double* src = fucoe.p + static_cast<int64_t>(same_flux ? 0 : d) * E() * P;
    partial_sums(tabs, flx, m, k, flux_term ? 0.5 : 1.0, src, rhs.p);
  }
}

void Context::rhs_single_dim(SweepTable tab, int d, double coef, const double* coef_inst) {
  int sz[kMaxDim];
  for (int i = 0; i < D; ++i) sz[i] = k + 1;
  sweep(u.p, sz, nullptr, rhs.p, sz, d, tab, true, LU_FULL, coef, k, k, coef_inst);
}

void Context::rhs_penalty(const double* alpha) {
  // per-instance constants replace alpha when the caller set them
  const bool per = !inst_alpha.empty();
  for (int d = 0; d < D; ++d)
    rhs_single_dim(table(8), d, -alpha[d] / 2., per ? d_pen_coef.p + static_cast<size_t>(d) * B() : nullptr);
}

void Context::upload_inst_alpha() {
  std::vector<double> pc(static_cast<size_t>(B()) * D);
  for (int i = 0; i < B(); ++i)
    for (int d = 0; d < D; ++d) pc[static_cast<size_t>(d) * B() + i] = -inst_alpha[static_cast<size_t>(i) * D + d] / 2.;
  upload(d_pen_coef, pc, stream);
}

void Context::rhs_linear(const double* coef) {
  if (lin_coef.size() != static_cast<size_t>(D) || !std::equal(lin_coef.begin(), lin_coef.end(), coef)) {
    OperatorTables aa(alpt, alpt);
    for (int d = 0; d < D; ++d) upload(t_lin[d], to_blocks(linear_operator_table(aa, coef[d]), 1 << nmax, k + 1, k + 1), stream);
    lin_coef.assign(coef, coef + D);
  }
  for (int d = 0; d < D; ++d) rhs_single_dim({t_lin[d].p, 1 << nmax}, d, 1.0);
}

void Context::rk_init() {
  if (E() > 0) hipLaunchKernelGGL(k_copy, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, tn.p, u.p, E() * K);
}

void Context::rk_stage(int scheme, int stage, double dt) {
  if (E() > 0) {
    prof_begin(PROF_RK, 8.0 * static_cast<double>(E()) * K * 4);
    hipLaunchKernelGGL(k_rk, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, u.p, tn.p, rhs.p, E() * K, scheme,
                       stage, dt, use_inst_dt ? d_inst_dt.p : nullptr, g_inst.p, K, err_flag.p);
    prof_end();
  }
}

void Context::copy_to_predict() {
  if (E() > 0) hipLaunchKernelGGL(k_copy, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, pred.p, u.p, E() * K);
}

void Context::copy_predict_to_ucoe() {
  if (E() > 0) hipLaunchKernelGGL(k_copy, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, u.p, pred.p, E() * K);
}

namespace {

void strides_of(const int* sz, int D, int* st) {
  st[D - 1] = 1;
  for (int d = D - 2; d >= 0; --d) st[d] = st[d + 1] * sz[d + 1];
}

int prod_of(const int* sz, int D) {
  int p = 1;
  for (int d = 0; d < D; ++d) p *= sz[d];
  return p;
}

}  // namespace

Context::FusedPlan& Context::fused_plan(const StageProblem& p) {
  // The terms hold the fiber-order arrays of the tile kernels: build them
  // first (they are reallocated only when the grid changes, which also
  // invalidates the plan).
  ensure_fiber_order();
  // one cached plan per buffer of the u / u_next pair (the stage swaps them)
  int slot = -1;
  for (int i = 0; i < 2; ++i)
    if (plans[i] && plans[i]->u == u.p) slot = i;
  const double* pen_coef = inst_alpha.empty() ? nullptr : d_pen_coef.p;
  if (slot >= 0) {
    FusedPlan& pl = *plans[slot];
    const bool same_alpha = std::equal(pl.alpha, pl.alpha + D, p.alpha);
    if (pl.version == grid_version && pl.tn == tn.p && pl.fucoe == fucoe.p && pl.unext == u_next.p &&
        pl.tmp == fused_tmp.p && pl.flux_kind == p.flux_kind && pl.interp == interp && same_alpha &&
        pl.precision == precision && pl.fucoe32 == fucoe32.p && pl.tv == tile_tv.p && pl.pen_coef == pen_coef &&
        pl.perm0 == g_perm[0].p && pl.pre0 == g_pre[0].p)
      return pl;
  } else {
    // keep the plan of the buffer the next stage starts from (u_next)
    slot = (plans[0] && plans[0]->u == u_next.p) ? 1 : 0;
  }
  not_recording("stage plan");
  HostTimer timer(*this, HOST_PLAN);
  // temporaries: (D-1) steps of 2^(D-1) eval_up chains and 2*D*2^(D-1) rhs chains
  std::vector<std::vector<int>> lus;
  const std::vector<std::vector<int>> dims = transform_orders(D, lus);
  const int nord = static_cast<int>(dims.size());
  const int mx = std::max(k, m) + 1;
  int T = 1;
  for (int d = 0; d < D; ++d) T *= mx;
  const int nchains = nord + 2 * D * nord;
  const size_t Ee = static_cast<size_t>(std::max<int64_t>(E(), 1));
  const size_t per = Ee * T;
  fused_tmp.ensure(per * nchains * std::max(D - 1, 1));
  u_next.ensure(Ee * K);
  // per-term value buffers of the fiber-tile last sweeps
  tile_tv.ensure(std::max(Ee * P * nord, Ee * K * 2 * D * nord));
  tile_pv.ensure(Ee * K * D);
  if (!plans[slot]) plans[slot].reset(new FusedPlan());
  FusedPlan& pl = *plans[slot];
  pl.version = grid_version;
  pl.interp = interp;
  pl.precision = precision;
  pl.fucoe32 = fucoe32.p;
  pl.pen_coef = pen_coef;
  pl.tv = tile_tv.p;
  pl.perm0 = g_perm[0].p;
  pl.pre0 = g_pre[0].p;
  const bool f32 = precision == 1;
  // algorithmic bytes of a term input: FP64 coefficients, else the storage word
  const double w = f32 ? 4.0 : 8.0;
  auto in_bytes = [&](const TermDev& t) { return (t.in32 || !f32) ? w : 8.0; };
  pl.u = u.p; pl.tn = tn.p; pl.fucoe = fucoe.p; pl.unext = u_next.p; pl.tmp = fused_tmp.p;
  pl.flux_kind = p.flux_kind;
  std::copy(p.alpha, p.alpha + D, pl.alpha);
  auto tmp = [&](int chain, int step) { return fused_tmp.p + (static_cast<size_t>(step) * nchains + chain) * per; };
  auto make_term = [&](const double* in, const float* in32, const int* sin_sz, int dd, SweepTable tab, int pt1,
                       bool flx, int lu, double coef, bool split) {
    TermDev t{};
    t.in = in; t.in32 = in32; t.M = tab.dev; t.ns = tab.ns; t.pt1 = pt1; t.dd = dd;
    if (split) {  // FP32 sweep: the table split in two FP32 parts
      const SplitTable& sp = split_table(tab.dev, static_cast<size_t>(tab.ns) * tab.ns * sin_sz[dd] * pt1);
      t.Mh = sp.hi.p;
      t.Ml = sp.lo.p;
    }
    t.slot = g_slot[dd].p; t.fbase = g_fbase[dd].p; t.fmaxl = g_fmaxl[dd].p; t.fmap = g_fmap[dd].p;
    t.fbits = g_fbits[dd].p;
    t.lsum = g_lsum.p;
    t.perm = g_perm[dd].p;
    t.pre = g_pre[dd].p;
    t.inst = g_inst.p;
    t.Tin = prod_of(sin_sz, D);
    std::copy(sin_sz, sin_sz + D, t.isz);
    strides_of(sin_sz, D, t.sin);
    t.flx = flx ? 1 : 0; t.lu = lu; t.coef = coef;
    return t;
  };
  // a term written to a buffer with output sizes osz
  auto make_task = [&](const TermDev& t, const int* osz, double* out) {
    TaskDev tk{};
    tk.t = t;
    tk.out = out;
    tk.out32 = reinterpret_cast<float*>(out);  // FP32 storage uses the first half of the buffer
    tk.Tout = prod_of(osz, D);
    std::copy(osz, osz + D, tk.osz);
    tk.R = t.Tin / t.isz[t.dd];
    return tk;
  };
  // Chains of partial sums (one per transform order and term), built as a
  // prefix tree: chains whose first sweeps have the same input, dimension,
  // table, relation and L/U mode share those sweeps' tasks and buffers.  A
  // chain's coefficient (1 or 1/2) is applied at the first sweep that only
  // chains with that coefficient pass through (the reference applies it in
  // the first sweep).  Scaling by a power of two commutes exactly with every
  // later product and sum (normal range), so every term value is
  // bit-identical to the unshared chain's.  Returns the last term of each
  // chain; appends the tasks of the shared sweeps per phase.
  struct ChainSpec {
    const double* src;
    const float* src32;
    SweepTable tabs[kMaxDim];
    int flx[kMaxDim];
    int o;
    double coef;
  };
  auto build_chains = [&](const std::vector<ChainSpec>& specs, int pf, int pt, int slot0,
                          std::vector<std::vector<TaskDev>>& phase_tasks) {
    auto prefix_key = [&](const ChainSpec& c, int st) {
      std::vector<int64_t> key = {reinterpret_cast<int64_t>(c.src)};
      for (int j = 0; j <= st; ++j) {
        const int dd = dims[c.o][j];
        key.insert(key.end(), {dd, reinterpret_cast<int64_t>(c.tabs[dd].dev), c.flx[dd], lus[c.o][j]});
      }
      return key;
    };
    std::map<std::vector<int64_t>, std::vector<double>> coefs;  // coefficients of the chains through a node
    for (const ChainSpec& c : specs)
      for (int st = 0; st + 1 < D; ++st) {
        std::vector<double>& v = coefs[prefix_key(c, st)];
        if (std::find(v.begin(), v.end(), c.coef) == v.end()) v.push_back(c.coef);
      }
    std::map<std::vector<int64_t>, TaskDev> nodes;
    std::vector<int> next_slot(std::max(D - 1, 1), slot0);
    std::vector<TermDev> terms;
    for (const ChainSpec& c : specs) {
      int sz[kMaxDim];
      for (int d = 0; d < D; ++d) sz[d] = pf + 1;
      const double* cur = c.src;
      const float* cur32 = c.src32;
      bool applied = false;
      for (int st = 0; st < D; ++st) {
        const int dd = dims[c.o][st];
        int sin_sz[kMaxDim];
        std::copy(sz, sz + D, sin_sz);
        sz[dd] = pt + 1;
        if (st == D - 1) {
          terms.push_back(make_term(cur, cur32, sin_sz, dd, c.tabs[dd], pt + 1, c.flx[dd] != 0, lus[c.o][st],
                                    applied ? 1.0 : c.coef, f32));
          break;
        }
        const std::vector<int64_t> key = prefix_key(c, st);
        const bool here = !applied && coefs[key].size() == 1;
        auto it = nodes.find(key);
        if (it == nodes.end()) {
          const TermDev t = make_term(cur, cur32, sin_sz, dd, c.tabs[dd], pt + 1, c.flx[dd] != 0, lus[c.o][st],
                                      here ? c.coef : 1.0, f32);
          it = nodes.emplace(key, make_task(t, sz, tmp(next_slot[st]++, st))).first;
          phase_tasks[st].push_back(it->second);
        }
        applied = applied || here;
        cur = it->second.out;
        cur32 = it->second.out32;
      }
    }
    return terms;
  };
  auto finish_phases = [&](std::vector<std::vector<TaskDev>>& ph, DevBuf<TaskDev>& dev, std::vector<int>& ntask,
                           std::vector<std::array<int64_t, 4>>& total, std::vector<double>& bytes,
                           std::vector<int>& rest) {
    std::vector<TaskDev> flat;
    rest.clear();
    ntask.clear();
    total.clear();
    bytes.clear();
    for (auto& tasks : ph) {
      std::array<int64_t, 4> begin{0, 0, 0, 0};
      double b = 0;
      for (TaskDev& tk : tasks) {
        for (int v = 0; v < 4; ++v) tk.begin_v[v] = begin[v];
        begin[0] += E() * tk.Tout;
        begin[1] += E() * tk.R;
        begin[2] += E() * ((tk.R + kRC - 1) / kRC);
        begin[3] += E() * ((tk.R + 1) / 2);
        b += E() * (in_bytes(tk.t) * tk.t.Tin + w * tk.Tout);
      }
      bytes.push_back(b);
      int R = 0;
      for (const TaskDev& tk : tasks) R = std::max(R, tk.R);
      rest.push_back(R);
      ntask.push_back(static_cast<int>(tasks.size()));
      total.push_back(begin);
      flat.insert(flat.end(), tasks.begin(), tasks.end());
    }
    upload(dev, flat, stream);
  };
  // eval_up chains
  {
    std::vector<std::vector<TaskDev>> ph(std::max(D - 1, 0));
    std::vector<TermDev> terms;
    std::vector<TaskDev> finals;
    SweepTable tabs[kMaxDim];
    int flx[kMaxDim];
    for (int d = 0; d < D; ++d) { tabs[d] = table(0); flx[d] = 0; }
    int psz[kMaxDim];
    for (int d = 0; d < D; ++d) psz[d] = m + 1;
    std::vector<ChainSpec> specs;
    for (int o = 0; o < nord; ++o) {
      ChainSpec c{u.p, nullptr, {}, {}, o, 1.0};
      std::copy(tabs, tabs + D, c.tabs);
      std::copy(flx, flx + D, c.flx);
      specs.push_back(c);
    }
    terms = build_chains(specs, k, m, 0, ph);
    for (int o = 0; o < nord; ++o)
      finals.push_back(make_task(terms[o], psz, tile_tv.p + static_cast<size_t>(o) * Ee * P));
    finish_phases(ph, pl.up_tasks, pl.up_phase_ntask, pl.up_phase_total, pl.up_phase_bytes, pl.up_phase_R);
    pl.up_final_R = finals[0].R;
    upload(pl.up_terms, terms, stream);
    upload(pl.up_final_tasks, finals, stream);
    pl.n_up_terms = static_cast<int>(terms.size());
    const int nc = (p.flux_kind == 0) ? 1 : D;
    double b = w * E() * P * (1 + nc) + (is_varcoef(p.flux_kind) ? 8.0 * E() * P * nc : 0.);
    for (const TermDev& t : terms) b += in_bytes(t) * E() * t.Tin;
    pl.up_final_bytes = b;
  }
  // rhs chains, in the reference order: volume terms for d = 0..D-1, then flux terms
  {
    std::vector<std::vector<TaskDev>> ph(std::max(D - 1, 0));
    std::vector<TermDev> terms;
    std::vector<TaskDev> finals;
    int ksz[kMaxDim];
    for (int d = 0; d < D; ++d) ksz[d] = k + 1;
    std::vector<ChainSpec> specs;
    for (int ft = 0; ft < 2; ++ft) {
      for (int d = 0; d < D; ++d) {
        ChainSpec c{};
        for (int dd = 0; dd < D; ++dd) {
          if (dd == d) { c.tabs[dd] = ft ? table(3) : table(2); c.flx[dd] = ft; }
          else { c.tabs[dd] = table(1); c.flx[dd] = 0; }
        }
        const int64_t coff = static_cast<int64_t>(same_flux ? 0 : d) * E() * P;
        c.src = fucoe.p + coff;
        c.src32 = f32 ? fucoe32.p + coff : nullptr;
        c.coef = ft ? 0.5 : 1.0;
        for (int o = 0; o < nord; ++o) {
          c.o = o;
          specs.push_back(c);
        }
      }
    }
    terms = build_chains(specs, m, k, nord, ph);
    for (size_t i = 0; i < terms.size(); ++i) finals.push_back(make_task(terms[i], ksz, tile_tv.p + i * Ee * K));
    finish_phases(ph, pl.rhs_tasks, pl.rhs_phase_ntask, pl.rhs_phase_total, pl.rhs_phase_bytes, pl.rhs_phase_R);
    pl.rhs_final_R = finals[0].R;
    upload(pl.rhs_terms, terms, stream);
    upload(pl.rhs_final_tasks, finals, stream);
    pl.n_rhs_terms = static_cast<int>(terms.size());
    double b = 8.0 * E() * K * (1 + 3 + D);  // rhs out, u, tn and u_next, penalty inputs
    for (const TermDev& t : terms) b += in_bytes(t) * E() * t.Tin;
    pl.rhs_final_bytes = b;
    std::vector<TermDev> pens;
    std::vector<TaskDev> pen_tasks;
    // the LF penalty stays FP64 in every precision; per-instance constants
    // (-alpha_d / 2 of each instance) when the caller set them
    for (int d = 0; d < D; ++d) {
      pens.push_back(make_term(u.p, nullptr, ksz, d, table(8), k + 1, true, LU_FULL, -p.alpha[d] / 2., false));
      if (pen_coef) pens.back().coef_inst = pen_coef + static_cast<size_t>(d) * B();
      pen_tasks.push_back(make_task(pens.back(), ksz, tile_pv.p + static_cast<size_t>(d) * Ee * K));
    }
    upload(pl.pens, pens, stream);
    upload(pl.pen_tasks, pen_tasks, stream);
    pl.pen_R = pen_tasks[0].R;
    pl.n_pen = D;
  }
  return pl;
}

// Fiber order of every dimension (perm, prefix popcounts) for the current grid.
void Context::ensure_fiber_order() {
  if (order_version == grid_version) return;
  not_recording("fiber order");
  HostTimer timer(*this, HOST_UPLOAD);
  for (int d = 0; d < D; ++d) {
    const int64_t nw = fbits_words[d];
    g_cnt.ensure(static_cast<size_t>(nw + 1));
    g_pre[d].ensure(static_cast<size_t>(nw + 1));
    g_perm[d].ensure(static_cast<size_t>(std::max<int64_t>(E(), 1)));
    const size_t sb = tile_scan_bytes(nw);
    scan_tmp.ensure(std::max<size_t>(sb, 1));
    tile_prefix(g_fbits[d].p, nw, g_cnt.p, g_pre[d].p, scan_tmp.p, scan_tmp.cap, stream);
  }
  order_version = grid_version;
}

// Tiles of whole fibers for the tile-size variant (te = 64 << which).
Context::TileSetH& Context::tile_set(int which) {
  ensure_fiber_order();
  TileSetH& ts = tsets[which];
  if (ts.version == grid_version) return ts;
  not_recording("tile set");
  HostTimer timer(*this, HOST_UPLOAD);
  ts.te = 64 << which;
  ts.ntile = std::max<int64_t>((E() + ts.te - 1) / ts.te, 1);
  ts.max_tile = ts.te - 1 + maxfib;
  for (int d = 0; d < D; ++d) {
    ts.tbeg[d].ensure(static_cast<size_t>(ts.ntile));
    ts.tend[d].ensure(static_cast<size_t>(ts.ntile));
    tile_order(g_fbits[d].p, g_pre[d].p, g_fbase[d].p, g_slot[d].p, E(), 1 << nmax, ts.te, ts.ntile, g_perm[d].p,
               ts.tbeg[d].p, ts.tend[d].p, stream);
  }
  ts.version = grid_version;
  return ts;
}

// One tile launch of ntask tasks with pf1 x pt1 blocks; throws (before any
// launch) when a tile of this size does not fit in LDS.
template <class T>
void Context::tile_launch(int v, int pf1, int pt1, const TaskDev* tasks, int R, int ntask, bool last) {
  if (order_version != grid_version) throw std::logic_error("tile launch: fiber order of an older grid");
  // variants 4..6: a thread per (target, rest index); 7..9: a thread per target
  const int which = (v - 4) % 3;
  const bool allr = v >= 7;
  TileSetH& ts = tile_set(which);
  if (tile_rest_chunk(pf1, R, ts.max_tile, sizeof(T)) < 1) throw std::runtime_error("tile: does not fit in LDS");
  const int32_t* tb[kMaxDim];
  const int32_t* te[kMaxDim];
  for (int d = 0; d < D; ++d) { tb[d] = ts.tbeg[d].p; te[d] = ts.tend[d].p; }
  tile_sweep<T>(pf1, pt1, tasks, ntask, ts.ntile, ts.max_tile, tb, te, D, lsum_cap, allr, last, stream);
}

void Context::stage_fused(const StageProblem& p, int scheme, int stage, double dt) {
  if (E() == 0) return;
  if (is_varcoef(p.flux_kind) &&
      (flux_kind != p.flux_kind || flux_param != p.flux_param || vcoef_version != grid_version)) {
    flux_kind = p.flux_kind;
    flux_param = p.flux_param;
    compute_vcoef();
  }
  flux_kind = p.flux_kind;
  flux_param = p.flux_param;
  same_flux = (p.flux_kind == 0);
  if (p.flux_kind == 1 && D != 2) throw std::runtime_error("sin/cos flux is two-dimensional");
  if (interp == 1) check_herm_flux(p.flux_kind, D);
  if (precision == 1) {
    const size_t e = static_cast<size_t>(std::max<int64_t>(E(), 1));
    if (up32.cap < e * P || fp32.cap < e * P * D) not_recording("FP32 buffers");
    up32.ensure(e * P);
    for (DevBuf<float>* b : {&fp32, &fucoe32, &interA32, &interB32}) b->ensure(e * P * D);
  }
  FusedPlan& pl = fused_plan(p);
  ++stage_count;
  if (precision == 1) stage_body<float>(p, pl, scheme, stage, dt);
  else stage_body<double>(p, pl, scheme, stage, dt);
}

// One fused stage with intermediates stored in T (double, or float with split
// tables).  Every kernel class runs one of the mappings: 0..3 per-element
// gathers (fused_elem.hip), 4..6 fiber tiles of 64, 128 or 256 elements
// (fused_tile.hip); all give identical values.
// The same stage as one persistent cooperative launch (fused_coop.hip).
template <class T>
void Context::stage_coop(const StageProblem& p, FusedPlan& pl, int scheme, int stage, double dt) {
  constexpr bool f32 = sizeof(T) == 4;
  if (interp != 0) throw std::runtime_error("persistent stage: the Lagrange family is needed");
  if (order_version != grid_version) throw std::logic_error("persistent stage: fiber order of an older grid");
  TileSetH& ts = tile_set(0);
  auto fits = [&](int pf1, int R, int tb) {
    if (tile_rest_chunk(pf1, R, ts.max_tile, tb) < 1) throw std::runtime_error("persistent stage: tile does not fit");
  };
  for (int R : pl.up_phase_R) fits(k + 1, R, sizeof(T));
  for (int R : pl.rhs_phase_R) fits(m + 1, R, sizeof(T));
  fits(k + 1, pl.up_final_R, sizeof(T));
  fits(m + 1, pl.rhs_final_R, sizeof(T));
  fits(k + 1, pl.pen_R, sizeof(double));
  CoopStage S{};
  const int ncomp = same_flux ? 1 : D;
  auto add = [&](int kind, int dim, const TaskDev* t1, int64_t n1, const TaskDev* t2, int64_t n2) {
    if (S.nphase >= kCoopMaxPhases) throw std::logic_error("persistent stage: too many phases");
    S.ph[S.nphase++] = CoopPhase{kind, dim, t1, t2, n1, n2};
  };
  int64_t off = 0;
  for (size_t s = 0; s < pl.up_phase_ntask.size(); ++s) {
    add(COOP_TILE_UP, 0, pl.up_tasks.p + off, pl.up_phase_ntask[s] * ts.ntile, nullptr, 0);
    off += pl.up_phase_ntask[s];
  }
  add(COOP_TILE_UP, 0, pl.up_final_tasks.p, pl.n_up_terms * ts.ntile, nullptr, 0);
  add(COOP_UP_COMBINE, 0, nullptr, E() * P, nullptr, 0);
  // surplus transform, dimension by dimension (Context::surplus_transform)
  const int p1 = m + 1;
  const SplitTable* w = f32 ? &split_table(t_wt.p, static_cast<size_t>(1 << nmax) * p1 * p1) : nullptr;
  int stride = P;
  for (int d = 0; d < D; ++d) {
    stride /= p1;
    SurplusArgs& a = S.sur[d];
    if constexpr (f32) {
      a.in32 = d == 0 ? fp32.p : (d % 2 == 1 ? interA32.p : interB32.p);
      a.out32 = d == D - 1 ? fucoe32.p : (d % 2 == 0 ? interA32.p : interB32.p);
      a.wh = w->hi.p;
      a.wl = w->lo.p;
    } else {
      a.in = d == 0 ? fp.p : (d % 2 == 1 ? interA.p : interB.p);
      a.out = d == D - 1 ? fucoe.p : (d % 2 == 0 ? interA.p : interB.p);
    }
    a.slot = g_slot[d].p; a.fbase = g_fbase[d].p; a.fmap = g_fmap[d].p;
    a.anc_slot = t_anc_slot.p; a.anc_point = t_anc_point.p; a.wt = t_wt.p; a.err = err_flag.p;
    a.E = E(); a.comp_stride = E() * P; a.P = P; a.p1 = p1; a.stride_d = stride; a.ncomp = ncomp;
    add(COOP_SURPLUS, d, nullptr, E() * P * ncomp, nullptr, 0);
  }
  off = 0;
  for (size_t s = 0; s < pl.rhs_phase_ntask.size(); ++s) {
    add(COOP_TILE_RHS, 0, pl.rhs_tasks.p + off, pl.rhs_phase_ntask[s] * ts.ntile, nullptr, 0);
    off += pl.rhs_phase_ntask[s];
  }
  add(COOP_TILE_RHS, 0, pl.rhs_final_tasks.p, pl.n_rhs_terms * ts.ntile, pl.pen_tasks.p, pl.n_pen * ts.ntile);
  add(COOP_RHS_COMBINE, 0, nullptr, E() * K, nullptr, 0);
  for (int d = 0; d < D; ++d) {
    S.ts.tbeg[d] = ts.tbeg[d].p;
    S.ts.tend[d] = ts.tend[d].p;
  }
  S.ntile = ts.ntile; S.max_tile = ts.max_tile; S.lds_bytes = kTileLds; S.D = D; S.cap = lsum_cap;
  S.tv = tile_tv.p;
  if constexpr (f32) { S.up = up32.p; S.fp = fp32.p; }
  else { S.up = up.p; S.fp = fp.p; }
  S.vcoef = vcoef.p; S.up_kind = p.flux_kind; S.ncomp = ncomp; S.n_up_terms = pl.n_up_terms; S.fs = p.flux_param;
  S.nP = E() * P;
  S.pv = tile_pv.p; S.n_rhs_terms = pl.n_rhs_terms; S.npen = pl.n_pen; S.K = K; S.nK = E() * K;
  S.rhs = rhs.p; S.u = u.p; S.tn = tn.p; S.unext = u_next.p; S.scheme = scheme; S.stage = stage; S.dt = dt;
  S.dtv = use_inst_dt ? d_inst_dt.p : nullptr; S.inst = g_inst.p; S.err = err_flag.p;
  prof_begin(PROF_RK, pl.up_final_bytes + pl.rhs_final_bytes);
  coop_stage<T>(k + 1, m + 1, S, device, stream);
  prof_end();
  ++stage_launches;
  std::swap(u.p, u_next.p);
  std::swap(u.cap, u_next.cap);
}

template <class T>
void Context::stage_body(const StageProblem& p, FusedPlan& pl, int scheme, int stage, double dt) {
  if (stage_launch == 1) {
    stage_coop<T>(p, pl, scheme, stage, dt);
    return;
  }
  // Hermite: the final eval_up kernel writes point values only (kind -1); the
  // point flux needs the element's other degrees of freedom and runs after it
  const int up_kind = interp == 1 ? -1 : p.flux_kind;
  constexpr bool f32 = sizeof(T) == 4;
  T* upp;
  T* fpp;
  if constexpr (f32) { upp = up32.p; fpp = fp32.p; }
  else { upp = up.p; fpp = fp.p; }
  const int tc = f32 ? 4 : 0;  // autotuner classes per precision
  const int ncomp = same_flux ? 1 : D;
  const double* dtv = use_inst_dt ? d_inst_dt.p : nullptr;
  T* tvT = reinterpret_cast<T*>(tile_tv.p);
  // point values and flux
  int64_t off = 0;
  for (size_t s = 0; s < pl.up_phase_ntask.size(); ++s) {
    prof_begin(PROF_SWEEP, pl.up_phase_bytes[s]);
    const TaskDev* tasks = pl.up_tasks.p + off;
    const int nt = pl.up_phase_ntask[s];
    const std::array<int64_t, 4> tot = pl.up_phase_total[s];
    tuned_launch(TUNE_UP_PHASE + tc, [&](int v) {
      if (v >= 4) tile_launch<T>(v, k + 1, m + 1, tasks, pl.up_phase_R[s], nt, false);
      else elem_phase<T>(k + 1, m + 1, tasks, nt, tot[v], D, lsum_cap, v, stream);
      ++stage_launches;
    });
    prof_end();
    off += pl.up_phase_ntask[s];
  }
  {
    prof_begin(PROF_FLUX, pl.up_final_bytes);
    tuned_launch(TUNE_UP_FINAL + tc, [&](int v) {
      if (v >= 4) {
        tile_launch<T>(v, k + 1, m + 1, pl.up_final_tasks.p, pl.up_final_R, pl.n_up_terms, true);
        up_combine<T>(tvT, pl.n_up_terms, E() * P, upp, fpp, vcoef.p, up_kind, p.flux_param, ncomp, stream);
        stage_launches += 2;
      } else {
        elem_up_final<T>(v, k + 1, m + 1, pl.up_terms.p, pl.n_up_terms, E(), P, D, upp, fpp, vcoef.p, up_kind,
                         p.flux_param, ncomp, lsum_cap, stream);
        ++stage_launches;
      }
    });
    if (interp == 1) {
      hipLaunchKernelGGL(k_flux_herm<T>, dim3(nblocks(E() * P)), dim3(kBlock), 0, stream, upp, fpp, E(), D,
                         p.flux_kind, T(p.flux_param), ncomp);
      ++stage_launches;
    }
    prof_end();
  }
  if constexpr (f32) surplus_transform32(fp32.p, fucoe32.p, interA32.p, interB32.p, ncomp);
  else surplus_transform(fp.p, fucoe.p, interA.p, interB.p, ncomp);
  stage_launches += D;
  off = 0;
  for (size_t s = 0; s < pl.rhs_phase_ntask.size(); ++s) {
    prof_begin(PROF_RHS_PHASE, pl.rhs_phase_bytes[s]);
    const TaskDev* tasks = pl.rhs_tasks.p + off;
    const int nt = pl.rhs_phase_ntask[s];
    const std::array<int64_t, 4> tot = pl.rhs_phase_total[s];
    tuned_launch(TUNE_RHS_PHASE + tc, [&](int v) {
      if (v >= 4) tile_launch<T>(v, m + 1, k + 1, tasks, pl.rhs_phase_R[s], nt, false);
      else elem_phase<T>(m + 1, k + 1, tasks, nt, tot[v], D, lsum_cap, v, stream);
      ++stage_launches;
    });
    prof_end();
    off += pl.rhs_phase_ntask[s];
  }
  {
    prof_begin(PROF_RK, pl.rhs_final_bytes);
    tuned_launch(TUNE_RHS_FINAL + tc, [&](int v) {
      if (v >= 4) {
        tile_launch<T>(v, m + 1, k + 1, pl.rhs_final_tasks.p, pl.rhs_final_R, pl.n_rhs_terms, true);
        tile_launch<double>(v, k + 1, k + 1, pl.pen_tasks.p, pl.pen_R, pl.n_pen, true);
        rhs_combine<T>(tvT, tile_pv.p, pl.n_rhs_terms, pl.n_pen, E(), K, rhs.p, u.p, tn.p, u_next.p, scheme, stage, dt,
                       dtv, g_inst.p, err_flag.p, stream);
        stage_launches += 3;
      } else {
        elem_rhs_final<T>(v, m + 1, k + 1, pl.rhs_terms.p, pl.n_rhs_terms, pl.pens.p, pl.n_pen, E(), K, D, rhs.p, u.p,
                          tn.p, u_next.p, scheme, stage, dt, dtv, g_inst.p, lsum_cap, err_flag.p, stream);
        ++stage_launches;
      }
    });
    prof_end();
  }
  std::swap(u.p, u_next.p);
  std::swap(u.cap, u_next.cap);
}

void Context::not_recording(const char* what) const {
  if (recording) throw NeedsBuild(what);
}

void Context::rk_step_eager(const StageProblem& p, int scheme, int nstage, double dt) {
  rk_init();
  for (int s = 0; s < nstage; ++s) stage_fused(p, scheme, s, dt);
}

// Everything a recorded step depends on besides the contents of device
// memory: versions of the grid and of the tables built from it, the device
// buffers (their addresses change when they grow), the problem, the launch
// settings and, without per-instance steps, dt.
// positions of the u / u_next addresses in a step key (a step swaps them)
constexpr size_t kKeyU = 3, kKeyUNext = 4;

std::vector<uintptr_t> Context::step_key(const StageProblem& p, int scheme, int nstage, double dt) const {
  auto bits = [](double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return static_cast<uintptr_t>(b);
  };
  auto ptr = [](const void* q) { return reinterpret_cast<uintptr_t>(q); };
  std::vector<uintptr_t> key = {
      static_cast<uintptr_t>(grid_version), static_cast<uintptr_t>(vcoef_version),
      static_cast<uintptr_t>(order_version), ptr(u.p), ptr(u_next.p), ptr(tn.p), ptr(rhs.p), ptr(up.p), ptr(fp.p),
      ptr(fucoe.p), ptr(interA.p), ptr(interB.p), ptr(fused_tmp.p), ptr(tile_tv.p), ptr(tile_pv.p), ptr(vcoef.p),
      ptr(err_flag.p), ptr(g_inst.p), ptr(d_pen_coef.p), ptr(up32.p), ptr(fp32.p), ptr(fucoe32.p), ptr(interA32.p),
      ptr(interB32.p), ptr(plans[0].get()), ptr(plans[1].get()), static_cast<uintptr_t>(precision),
      static_cast<uintptr_t>(interp), static_cast<uintptr_t>(stage_launch), static_cast<uintptr_t>(kernel_variant + 1),
      static_cast<uintptr_t>(lsum_cap), static_cast<uintptr_t>(E()), tuned.size(), static_cast<uintptr_t>(scheme),
      static_cast<uintptr_t>(nstage), static_cast<uintptr_t>(p.flux_kind), bits(p.flux_param),
      static_cast<uintptr_t>(use_inst_dt), use_inst_dt ? ptr(d_inst_dt.p) : bits(dt), inst_alpha.size()};
  for (int d = 0; d < D; ++d) key.push_back(bits(p.alpha[d]));
  if (key[kKeyU] != ptr(u.p) || key[kKeyUNext] != ptr(u_next.p)) throw std::logic_error("step key layout");
  return key;
}

void Context::drop_step_graphs() {
  bool any = false;
  for (StepGraph& g : step_graphs) any = any || g.exec;
  if (any) sync();  // no recorded step is still running when it is destroyed
  for (StepGraph& g : step_graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.key.clear();
  }
}

void Context::rk_step(const StageProblem& p, int scheme, int nstage, double dt) {
  if (!graph_step || prof_on || E() == 0 || nstage < 1) {
    rk_step_eager(p, scheme, nstage, dt);
    return;
  }
  const std::vector<uintptr_t> key = step_key(p, scheme, nstage, dt);
  // the key without the u / u_next pair, which every odd step swaps
  auto norm = [](std::vector<uintptr_t> k) {
    k[kKeyU] = k[kKeyUNext] = 0;
    return k;
  };
  auto swaps = [&]() {
    for (int s = 0; s < nstage; ++s) {
      std::swap(u.p, u_next.p);
      std::swap(u.cap, u_next.cap);
    }
  };
  for (StepGraph& g : step_graphs) {
    if (g.exec && g.key == key) {
      hip_check(hipGraphLaunch(g.exec, stream), "hipGraphLaunch");
      ++graph_replays;
      stage_count += nstage;
      stage_launches += g.launches;
      swaps();
      return;
    }
  }
  // A step is recorded only when the same step already ran without a graph
  // on the same state: that run builds the plans, fiber orders and point-flux
  // tables and tunes the kernels, which upload host data and time launches
  // (nothing a graph can hold).
  const std::vector<uintptr_t> nk = norm(key);
  if (std::find(step_seen.begin(), step_seen.end(), nk) == step_seen.end()) {
    rk_step_eager(p, scheme, nstage, dt);
    if (step_seen.size() >= kStepGraphs) step_seen.erase(step_seen.begin());
    step_seen.push_back(norm(step_key(p, scheme, nstage, dt)));
    return;
  }
  StepGraph& g = step_graphs[next_graph];
  next_graph = (next_graph + 1) % kStepGraphs;
  if (g.exec) {
    sync();
    (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
  }
  g.key.clear();
  const int64_t launches0 = stage_launches, stages0 = stage_count;
  double* const u0 = u.p;
  double* const un0 = u_next.p;
  const size_t ucap0 = u.cap, uncap0 = u_next.cap;
  hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  hipGraph_t graph = nullptr;
  recording = true;
  try {
    rk_step_eager(p, scheme, nstage, dt);
  } catch (const std::exception& ex) {
    // drop the recording and undo the stages' host-side swaps and counters
    // (nothing of the step ran); when something had to be built first
    // (NeedsBuild: it threw before changing anything), run the step
    // without a graph
    recording = false;
    (void)hipStreamEndCapture(stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    u.p = u0; u_next.p = un0; u.cap = ucap0; u_next.cap = uncap0;
    stage_launches = launches0;
    stage_count = stages0;
    if (!dynamic_cast<const NeedsBuild*>(&ex)) throw;
    rk_step_eager(p, scheme, nstage, dt);
    return;
  }
  recording = false;
  hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
  size_t nodes = 0;
  hipError_t e = hipGraphGetNodes(graph, nullptr, &nodes);
  // a recorded persistent step is one copy and one launch per stage: a
  // stage the runtime did not record must not be replayed
  if (e == hipSuccess && stage_launch == 1 && nodes != static_cast<size_t>(1 + stage_launches - launches0))
    e = hipErrorStreamCaptureUnsupported;
  if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    g.exec = nullptr;
    throw std::runtime_error(std::string("step graph: recording failed: ") + hipGetErrorString(e));
  }
  g.launches = stage_launches - launches0;
  g.key = key;
  ++graph_captures;
  // recording ran nothing on the device: run the step now (its host-side
  // buffer swaps and counters were done while recording)
  hip_check(hipGraphLaunch(g.exec, stream), "hipGraphLaunch");
}

void Context::save_state(int slot) {
  if (slot < 0 || slot >= 4) throw std::runtime_error("snapshot slot must be 0..3");
  ensure_host_grid();
  Snapshot& sn = snaps[slot];
  sn.slots.clear();
  for (const HostGrid& g : grids) sn.slots.push_back(g.slot);
  sn.E = E();
  sn.u.ensure(std::max<int64_t>(E() * K, 1));
  if (E() > 0) hipLaunchKernelGGL(k_copy, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, sn.u.p, u.p, E() * K);
}

void Context::restore_state(int slot) {
  if (slot < 0 || slot >= 4) throw std::runtime_error("snapshot slot must be 0..3");
  Snapshot& sn = snaps[slot];
  if (sn.slots.empty()) throw std::runtime_error("snapshot slot is empty");
  // (with a stale host copy the grid is rebuilt on the device without comparing)
  bool same = !host_stale && static_cast<int>(sn.slots.size()) == B();
  for (int i = 0; same && i < B(); ++i) same = (sn.slots[i] == grids[i].slot);
  if (!same) {
    if (device_adapt && static_cast<int>(sn.slots.size()) == B()) set_grids_fresh(sn.slots);
    else set_grids(sn.slots, false);
  }
  if (E() > 0) hipLaunchKernelGGL(k_copy, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, u.p, sn.u.p, E() * K);
}

void Context::check_error_flag() {
  int32_t h = 0;
  hip_check(hipMemcpyAsync(&h, err_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, stream), "download");
  sync();
  if (h != 0) {
    int32_t z = 0;
    hip_check(hipMemcpyAsync(err_flag.p, &z, sizeof(int32_t), hipMemcpyHostToDevice, stream), "upload");
    sync();
    if (h & ERR_ANCESTOR) throw std::runtime_error("surplus transform: missing ancestor element (grid has a hole)");
    if (h & ERR_COOP) throw std::runtime_error("persistent stage: kernel was not launched cooperatively");
    throw std::runtime_error("RK stage: non-finite value in the solution");
  }
}

void Context::indicator_sumsq_dev() {
  if (E() > 0) hipLaunchKernelGGL(k_sumsq, dim3(nblocks(E())), dim3(kBlock), 0, stream, u.p, sumsq.p, E(), K);
}

void Context::indicator_sumsq(double* host_out) {
  if (E() == 0) return;
  indicator_sumsq_dev();
  hip_check(hipMemcpyAsync(host_out, sumsq.p, E() * sizeof(double), hipMemcpyDeviceToHost, stream), "download");
  sync();
}

int64_t Context::refine(double eps) { return refine_instances(std::vector<double>(B(), eps)); }

int64_t Context::coarsen(double eta) { return coarsen_instances(std::vector<double>(B(), eta)); }

int64_t Context::refine_instances(const std::vector<double>& eps) {
  HostTimer timer(*this, HOST_ADAPT);
  if (static_cast<int>(eps.size()) != B()) throw std::runtime_error("refine: one threshold per instance expected");
  if (device_adapt) {
    const int64_t r = refine_device(eps);
    if (r >= 0) return r;
  }
  ensure_host_grid();
  std::vector<double> s(E());
  indicator_sumsq(s.data());
  std::vector<std::vector<int32_t>> ns(B());
  std::vector<int64_t> added(B(), 0);
  parallel_for(B(), [&](int i) {
    std::vector<double> ni(off[i + 1] - off[i]);
    for (size_t e = 0; e < ni.size(); ++e) ni[e] = threshold_norm(s[off[i] + e], eps[i]);
    ns[i] = refine_slots(grids[i], ni, eps[i], added[i]);
  });
  int64_t tot = 0;
  std::vector<char> changed(B(), 0);
  for (int i = 0; i < B(); ++i) { tot += added[i]; changed[i] = added[i] > 0; }
  if (tot > 0) set_grids(ns, true, &changed);
  return tot;
}

int64_t Context::coarsen_instances(const std::vector<double>& eta) {
  HostTimer timer(*this, HOST_ADAPT);
  if (static_cast<int>(eta.size()) != B()) throw std::runtime_error("coarsen: one threshold per instance expected");
  if (device_adapt) {
    const int64_t r = coarsen_device(eta);
    if (r >= 0) return r;
  }
  ensure_host_grid();
  std::vector<double> s(E());
  indicator_sumsq(s.data());
  std::vector<std::vector<int32_t>> ns(B());
  std::vector<int64_t> removed(B(), 0);
  parallel_for(B(), [&](int i) {
    std::vector<double> ni(off[i + 1] - off[i]);
    for (size_t e = 0; e < ni.size(); ++e) ni[e] = threshold_norm(s[off[i] + e], eta[i]);
    ns[i] = coarsen_slots(grids[i], ni, eta[i], removed[i]);
  });
  int64_t tot = 0;
  std::vector<char> changed(B(), 0);
  for (int i = 0; i < B(); ++i) { tot += removed[i]; changed[i] = removed[i] > 0; }
  if (tot > 0) set_grids(ns, true, &changed);
  return tot;
}

void Context::init_adaptive_intp(double (*f)(const double*, int, void*), void* user, double eps) {
  std::vector<void*> users(B(), user);
  init_adaptive_intp_batch(f, users.data(), eps);
}

void Context::init_adaptive_intp_batch(double (*f)(const double*, int, void*), void* const* users, double eps) {
  ensure_host_grid();
  // initialisation is never restricted to a coarse grid
  struct CapGuard {
    int& c; int saved;
    explicit CapGuard(int& cap) : c(cap), saved(cap) { c = -1; }
    ~CapGuard() { c = saved; }
  } cap_guard(lsum_cap);
  LagrangeScope lagrange_scope(*this);
  if (up_host.size() != static_cast<size_t>(E()) * P) {
    up_host.assign(static_cast<size_t>(E()) * P, 0.);
    new_add.assign(E(), 1);
  }
  auto eval_new = [&]() {
    std::vector<double> x(D);
    for (int64_t e = 0; e < E(); ++e) {
      if (!new_add[e]) continue;
      void* user = users[uinst[e]];
      for (int q = 0; q < P; ++q) {
        int rem = q;
        int idx[kMaxDim];
        for (int d = D - 1; d >= 0; --d) { idx[d] = rem % (m + 1); rem /= (m + 1); }
        for (int d = 0; d < D; ++d) x[d] = lagr.point(uslot[e * D + d] * (m + 1) + idx[d]);
        up_host[e * P + q] = f(x.data(), D, user);
      }
    }
    upload(up, up_host, stream);
    ensure_fields();
    surplus_transform(up.p, uintp.p, interA.p, interB.p, 1);
  };
  eval_new();
  while (true) {
    std::fill(new_add.begin(), new_add.end(), 0);
    LagrNormArgs a{};
    a.c = uintp.p; a.lmass = t_lmass.p; a.out = sumsq.p; a.E = E(); a.D = D; a.P = P; a.p1 = m + 1;
    for (int d = 0; d < D; ++d) a.slot[d] = g_slot[d].p;
    if (E() > 0) hipLaunchKernelGGL(k_lagr_norm, dim3(nblocks(E())), dim3(kBlock), 0, stream, a);
    std::vector<double> s(E());
    if (E() > 0)
      hip_check(hipMemcpyAsync(s.data(), sumsq.p, E() * sizeof(double), hipMemcpyDeviceToHost, stream), "download");
    sync();
    check_error_flag();
    for (double& v : s) v = std::pow(v, 0.5);
    std::vector<std::vector<int32_t>> ns(B());
    int64_t added = 0;
    for (int i = 0; i < B(); ++i) {
      const std::vector<double> ni(s.begin() + off[i], s.begin() + off[i + 1]);
      int64_t ai = 0;
      // per-instance thresholds, when the caller set them, also refine the initialisation
      const double ei = static_cast<int>(inst_eps.size()) == B() ? inst_eps[i] : eps;
      ns[i] = refine_slots(grids[i], ni, ei, ai);
      added += ai;
    }
    if (added == 0) break;
    set_grids(ns, true);
    eval_new();
  }
  // FastLagrInit::eval_ucoe_Alpt_Lagr: Lagrange surplus -> Alpert coefficients
  if (E() > 0) hipLaunchKernelGGL(k_fill, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, u.p, E() * K, 0.);
  SweepTable tabs[kMaxDim];
  int flx[kMaxDim];
  for (int d = 0; d < D; ++d) { tabs[d] = table(1); flx[d] = 0; }
  partial_sums(tabs, flx, m, k, 1.0, uintp.p, u.p);
  sync();
  // the host point values only serve the incremental refinement above; keeping
  // them would make every later grid change copy E*P values on the host
  std::vector<double>().swap(up_host);
  std::vector<char>().swap(new_add);
}

// Values at the interpolation points (degrees of freedom) of the current
// family -> Alpert coefficients: hierarchical surplus transform, then the
// interpolation-to-Alpert partial sums (FastLagrInit::eval_ucoe_Alpt_Lagr,
// source/FastMultiplyLU.cpp:1617-1620; for the Hermite family
// HermInterpolation::init_from_Herm_to_alpt, source/Interplation.cpp:5019-5030).
void Context::init_from_points(const double* values) {
  ensure_host_grid();
  std::vector<double> h(values, values + static_cast<size_t>(E()) * P);
  upload(up, h, stream);
  ensure_fields();
  surplus_transform(up.p, uintp.p, interA.p, interB.p, 1);
  if (E() > 0) hipLaunchKernelGGL(k_fill, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, u.p, E() * K, 0.);
  SweepTable tabs[kMaxDim];
  int flx[kMaxDim];
  for (int d = 0; d < D; ++d) { tabs[d] = table(1); flx[d] = 0; }
  partial_sums(tabs, flx, m, k, 1.0, uintp.p, u.p);
  check_error_flag();
}

// fp = coef * du/dx at the points with weight w > 0: w counts the viscosity
// supports holding the point (overlap) or marks it (w = 1)
__global__ __launch_bounds__(kBlock) void k_visc_points(const double* up, const double* w, double* fp, int64_t n,
                                                         double coef, int overlap) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double wi = w[i];
  fp[i] = wi == 0. ? 0. : (overlap ? wi * coef * up[i] : coef * up[i]);
}

void Context::eval_der_up(int d0) {
  if (interp != 0) throw std::runtime_error("derivative point values: the Lagrange family is needed");
  if (t_pt_d1.cap == 0)
    upload(t_pt_d1, to_blocks(alpt_at_lagr_points(alpt, lagr, 1), 1 << nmax, k + 1, m + 1), stream);
  if (E() > 0) hipLaunchKernelGGL(k_fill, dim3(nblocks(E() * P)), dim3(kBlock), 0, stream, up.p, E() * P, 0.);
  SweepTable tabs[kMaxDim];
  int flx[kMaxDim];
  for (int d = 0; d < D; ++d) { tabs[d] = table(0); flx[d] = 0; }
  tabs[d0] = {t_pt_d1.p, 1 << nmax};
  partial_sums(tabs, flx, k, m, 1.0, u.p, up.p);
}

int64_t Context::viscosity_elements(double kappa) {
  ensure_host_grid();
  std::vector<double> s(E());
  indicator_sumsq(s.data());
  visc_elems.clear();
  visc_keys.assign(B(), {});
  for (int i = 0; i < B(); ++i) {
    const HostGrid& g = grids[i];
    for (int64_t e = 0; e < g.E; ++e) {
      int32_t sl[kMaxDim];
      int lmax = 0, lsum = 0, have = 0;
      for (int d = 0; d < D; ++d) {
        const int l = level_of_slot(g.slot[e * D + d]);
        lmax = std::max(lmax, l);
        lsum += l;
      }
      if (lmax != nmax) continue;
      for (int d = 0; d < D && have == 0; ++d) {
        for (int q = 0; q < D; ++q) sl[q] = g.slot[e * D + q];
        const int sd = sl[d], l = level_of_slot(sd);
        if (l == 0) { sl[d] = 1; have += g.find(sl) >= 0; }
        else if (l < nmax) {
          sl[d] = 2 * sd; have += g.find(sl) >= 0;
          sl[d] = 2 * sd + 1; have += g.find(sl) >= 0;
        }
      }
      if (have != 0) continue;
      const double num_l2_norm = std::pow(s[off[i] + e], 1. / 2.);
      const double theory_bound = std::pow(2., (-k + 1.0) * lsum);
      if (std::log10(num_l2_norm) - std::log10(theory_bound) > kappa) {
        visc_elems.push_back(off[i] + e);
        visc_keys[i].push_back(HostGrid::pack(&g.slot[e * D], D));
      }
    }
  }
  return static_cast<int64_t>(visc_elems.size());
}

void Context::viscosity_flux(double coef, bool overlap) {
  ensure_host_grid();
  LagrangeScope lagrange_scope(*this);
  const int p1 = m + 1;
  std::vector<double> w(static_cast<size_t>(E()) * P, 0.);
  // supports of the viscosity elements, per instance
  std::vector<std::vector<double>> box(B());
  for (int i = 0; i < B() && i < static_cast<int>(visc_keys.size()); ++i)
    for (uint64_t key : visc_keys[i]) {
      const int32_t le = grids[i].index.find(key);
      if (le < 0) continue;  // removed since the update
      const int64_t v = off[i] + le;
      for (int d = 0; d < D; ++d) {
        double xl, xr;
        cell_support(level(v, d), suppt_of_slot(uslot[v * D + d]), xl, xr);
        box[i].push_back(xl);
        box[i].push_back(xr);
      }
    }
  parallel_for(B(), [&](int i) {
    const std::vector<double>& bx = box[i];
    const int nv = static_cast<int>(bx.size() / (2 * D));
    if (nv == 0) return;
    for (int64_t e = off[i]; e < off[i + 1]; ++e) {
      // DGSolution::update_viscosity_intersect_element: supports meet in every dimension
      double exl[kMaxDim], exr[kMaxDim];
      for (int d = 0; d < D; ++d) cell_support(level(e, d), suppt_of_slot(uslot[e * D + d]), exl[d], exr[d]);
      bool any = false;
      for (int v = 0; v < nv && !any; ++v) {
        bool meet = true;
        for (int d = 0; d < D && meet; ++d) {
          const double vl = bx[(v * D + d) * 2], vr = bx[(v * D + d) * 2 + 1];
          meet = !((exl[d] >= vr) || (exr[d] <= vl));
        }
        any = meet;
      }
      if (!any) continue;
      for (int q = 0; q < P; ++q) {
        double pos[kMaxDim];
        int rem = q;
        for (int d = D - 1; d >= 0; --d) {
          pos[d] = lagr.point(uslot[e * D + d] * p1 + rem % p1);
          rem /= p1;
        }
        int num = 0;
        for (int v = 0; v < nv; ++v) {
          bool in = true;
          for (int d = 0; d < D && in; ++d) in = pos[d] >= bx[(v * D + d) * 2] && pos[d] <= bx[(v * D + d) * 2 + 1];
          if (in) { ++num; if (!overlap) break; }
        }
        w[e * P + q] = num;
      }
    }
  });
  upload(visc_w, w, stream);
  const int64_t n = E() * P;
  for (int d0 = 0; d0 < D; ++d0) {
    eval_der_up(d0);
    if (n > 0)
      hipLaunchKernelGGL(k_visc_points, dim3(nblocks(n)), dim3(kBlock), 0, stream, up.p, visc_w.p, fp.p + d0 * n, n,
                         coef, overlap ? 1 : 0);
  }
  same_flux = false;
  surplus_transform(fp.p, fucoe.p, interA.p, interB.p, D);
}

// DGAdapt::init_separable_scalar_sum (source/DGAdapt.cpp:120-147): project
// sum_i prod_d f_i(x_d) onto the grid, then refine every leaf whose
// coefficient norm exceeds eps with separably initialised children
// (refine_init_separable_scalar_sum, :850-894) until no leaf does.
void Context::init_separable(double (*f)(double, int, int, void*), void* user, int nfunc, double eps) {
  ensure_host_grid();
  if (B() != 1) throw std::runtime_error("separable initialisation: one instance expected");
  if (nfunc < 1) throw std::runtime_error("separable initialisation: at least one function expected");
  // DGSolution::seperable_project (source/DGSolution.cpp:150-167)
  std::vector<std::vector<double>> c1(static_cast<size_t>(nfunc) * D);
  for (int i = 0; i < nfunc; ++i)
    for (int d = 0; d < D; ++d) c1[i * D + d] = project_1d(alpt, [&](double x) { return f(x, d, i, user); });
  // DGSolution::init_elem_separable (source/DGSolution.cpp:169-184), functions
  // in order, added to the coefficients the elements hold (new ones hold zero)
  const HostGrid g0 = grids[0];
  std::vector<double> u0(static_cast<size_t>(g0.E) * K);
  if (g0.E > 0)
    hip_check(hipMemcpyAsync(u0.data(), u.p, u0.size() * sizeof(double), hipMemcpyDeviceToHost, stream), "download");
  sync();
  std::vector<double> uh;
  auto coefficients = [&](const HostGrid& g) {
    const std::vector<int32_t> old = map_new_to_old(g0, g);
    uh.assign(static_cast<size_t>(g.E) * K, 0.);
    for (int64_t e = 0; e < g.E; ++e)
      if (old[e] >= 0) std::copy(&u0[static_cast<size_t>(old[e]) * K], &u0[static_cast<size_t>(old[e]) * K] + K, &uh[e * K]);
    for (int64_t e = 0; e < g.E; ++e)
      for (int q = 0; q < K; ++q) {
        int idx[kMaxDim];
        int rem = q;
        for (int d = D - 1; d >= 0; --d) { idx[d] = rem % (k + 1); rem /= (k + 1); }
        for (int i = 0; i < nfunc; ++i) {
          double coe = 1.;
          for (int d = 0; d < D; ++d) coe *= c1[i * D + d][g.slot[e * D + d] * (k + 1) + idx[d]];
          uh[e * K + q] += coe;
        }
      }
  };
  while (true) {
    coefficients(grids[0]);
    std::vector<double> norm(grids[0].E);
    for (int64_t e = 0; e < grids[0].E; ++e) {  // VecMultiD::norm(2.)
      double sq = 0.;
      for (int q = 0; q < K; ++q) sq += std::pow(uh[e * K + q], 2.);
      norm[e] = std::pow(sq, 1. / 2.);
    }
    int64_t added = 0;
    std::vector<int32_t> ns = refine_slots(grids[0], norm, eps, added);
    if (added == 0) break;
    set_grid_slots(ns, false);
  }
  set_grid_slots(grids[0].slot, false);
  upload(u, uh, stream);
  sync();
}

// DGAdapt::filter / filter_local (source/DGAdapt.cpp:461-555): every element
// with level sum >= start is scaled by exp(-c (sum_d a_d 2^l_d)^2),
// c = damp * dt^2/2 * pi^2; speed[e*D+d] is a_d for element e.
void Context::filter(double damp, const double* speed, double dt, int start) {
  ensure_host_grid();
  const double coefficient = damp * (dt * dt / 2.0) * (M_PI * M_PI);
  std::vector<double> fac(E(), 1.);
  std::vector<char> on(E(), 0);
  for (int64_t e = 0; e < E(); ++e) {
    int ls = 0;
    for (int d = 0; d < D; ++d) ls += level(e, d);
    if (ls < start) continue;
    double index_sum = 0.0;
    for (int d = 0; d < D; ++d) index_sum += speed[e * D + d] * std::pow(2.0, level(e, d));
    fac[e] = std::exp(-coefficient * index_sum * index_sum);
    on[e] = 1;
  }
  std::vector<double> f(static_cast<size_t>(E()) * K);
  for (int64_t e = 0; e < E(); ++e)
    for (int q = 0; q < K; ++q) f[e * K + q] = on[e] ? fac[e] : 1.;
  ensure_fields();  // tmpA holds at least E*K values, so the upload below keeps it
  upload(tmpA, f, stream);
  if (E() > 0) hipLaunchKernelGGL(k_scale, dim3(nblocks(E() * K)), dim3(kBlock), 0, stream, u.p, tmpA.p, E() * K);
}

void Context::eval_tensor(int nx, const double* x1d, double* out_host) {
  ensure_host_grid();
  if (B() != 1) throw std::runtime_error("point evaluation: one instance expected");
  if (nx < 1) throw std::runtime_error("point evaluation: empty coordinate set");
  int64_t npts = 1;
  for (int d = 0; d < D; ++d) {
    npts *= nx;
    if (npts > (int64_t{1} << 34)) throw std::runtime_error("point evaluation: grid too large");
  }
  // slot -> element table of the current grid
  if (eval_version != grid_version) {
    uint64_t cap = 2;
    while (cap < 2 * static_cast<uint64_t>(std::max<int64_t>(E(), 1))) cap <<= 1;
    std::vector<uint64_t> hk_(cap, kEmptyKey);
    std::vector<int32_t> hv(cap, -1);
    for (int64_t e = 0; e < E(); ++e) {
      uint64_t key = 0;
      for (int d = 0; d < D; ++d) key |= static_cast<uint64_t>(uslot[e * D + d]) << (16 * d);
      uint64_t h = mix64(key) & (cap - 1);
      while (hk_[h] != kEmptyKey) h = (h + 1) & (cap - 1);
      hk_[h] = key;
      hv[h] = static_cast<int32_t>(e);
    }
    upload(eval_hkey, hk_, stream);
    upload(eval_hval, hv, stream);
    eval_hmask = cap - 1;
    eval_version = grid_version;
  }
  // cells whose closed support holds each coordinate, with the basis values there
  const int k1 = k + 1, cm = 2 * (nmax + 1);
  std::vector<int32_t> ncand(nx, 0), cslot(static_cast<size_t>(nx) * cm, 0);
  std::vector<double> cval(static_cast<size_t>(nx) * cm * k1, 0.);
  for (int i = 0; i < nx; ++i) {
    const double x = x1d[i];
    int n = 0;
    auto add = [&](int level, int suppt) {
      double xl, xr;
      cell_support(level, suppt, xl, xr);
      if (x < xl || x > xr) return;
      const int sl = slot_of(level, suppt);
      cslot[i * cm + n] = sl;
      for (int p = 0; p < k1; ++p) cval[(static_cast<size_t>(i) * cm + n) * k1 + p] = alpt.val(sl * k1 + p, x, 0, 0);
      ++n;
    };
    for (int l = 0; l <= nmax; ++l) {
      if (l <= 1) { add(l, 1); continue; }
      const int ncell = 1 << (l - 1);
      const int c0 = static_cast<int>(std::floor(x * ncell));
      for (int cc = c0 - 1; cc <= c0 + 1; ++cc)
        if (cc >= 0 && cc < ncell) add(l, 2 * cc + 1);
    }
    ncand[i] = n;
  }
  upload(eval_ncand, ncand, stream);
  upload(eval_cslot, cslot, stream);
  upload(eval_cval, cval, stream);
  eval_out.ensure(static_cast<size_t>(npts));
  EvalArgs a{};
  a.u = u.p; a.hkey = eval_hkey.p; a.hval = eval_hval.p; a.hmask = eval_hmask;
  a.ncand = eval_ncand.p; a.cslot = eval_cslot.p; a.cval = eval_cval.p; a.out = eval_out.p;
  a.npts = npts; a.nx = nx; a.cm = cm; a.k1 = k1; a.D = D; a.K = K;
  hipLaunchKernelGGL(k_eval_tensor, dim3(nblocks(npts)), dim3(kBlock), 0, stream, a);
  hip_check(hipMemcpyAsync(out_host, eval_out.p, npts * sizeof(double), hipMemcpyDeviceToHost, stream), "download");
  sync();
}

// Thread-mapping autotuner of the fused stage kernels.  The four mappings
// compute identical values (same sums in the same order), so the first time
// a kernel class meets a grid-size class, every mapping is launched once on
// the live data with events around it and the fastest is kept; the outputs of
// the last launch stand.  A forced mapping (kernel_variant >= 0) skips this.
// A mapping whose block does not fit (its shape check throws before any
// launch) is skipped; mapping 0 always fits.
namespace {
// Decisions of the autotuner, shared by the contexts of the process (keyed
// by everything that shapes the kernels' work besides the grid itself).
std::mutex g_tuned_mutex;
std::map<std::vector<int64_t>, int> g_tuned;
}  // namespace

void Context::tuned_launch(int cls, const std::function<void(int)>& launch) {
  auto try_launch = [&](int v) {
    try {
      launch(v);
      return true;
    } catch (const std::runtime_error&) {
      if (v == 0) throw;
      return false;
    }
  };
  const int fixed = kernel_variant >= 0 ? kernel_variant : class_variant[cls];
  if (fixed >= 0) {
    if (!try_launch(fixed)) launch(0);
    return;
  }
  int bucket = 0;
  while ((int64_t{1} << (bucket + 1)) <= E()) ++bucket;
  const int key = cls * 64 + bucket;
  auto it = tuned.find(key);
  if (it != tuned.end()) {
    launch(it->second);
    return;
  }
  const std::vector<int64_t> gkey = {D, k, m, interp, precision, lsum_cap >= 0, cls, bucket};
  {
    std::lock_guard<std::mutex> lock(g_tuned_mutex);
    auto g = g_tuned.find(gkey);
    if (g != g_tuned.end()) {
      tuned[key] = g->second;
      launch(g->second);
      return;
    }
  }
  not_recording("kernel tuning");
  // every mapping runs once untimed (first-launch loading, lazily built
  // tile sets), then its time is the least of three timed runs; the stage
  // kernels do not read their own outputs, so repeating a launch changes
  // no value
  struct Events {  // released however the loop ends
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  hip_check(hipEventCreate(&ev.a), "hipEventCreate");
  hip_check(hipEventCreate(&ev.b), "hipEventCreate");
  float best = 0.f;
  int bv = 0;
  for (int v = 0; v < kVariants; ++v) {
    if (!try_launch(v)) continue;
    float tv = 0.f;
    for (int r = 0; r < 3; ++r) {
      hip_check(hipEventRecord(ev.a, stream), "hipEventRecord");
      launch(v);
      hip_check(hipEventRecord(ev.b, stream), "hipEventRecord");
      hip_check(hipEventSynchronize(ev.b), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, ev.a, ev.b), "hipEventElapsedTime");
      tv = r == 0 ? ms : std::min(tv, ms);
    }
    if (v == 0 || tv < best) { best = tv; bv = v; }
  }
  tuned[key] = bv;
  std::lock_guard<std::mutex> lock(g_tuned_mutex);
  g_tuned.emplace(gkey, bv);
}

}  // namespace adg
