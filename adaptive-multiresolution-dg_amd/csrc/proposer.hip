// Learned refinement proposer: element features, the MLP's training and
// inference kernels, and the adaptive step that uses it.
//
// The reference refines by predicting u^{n+1} with one forward-Euler stage
// and giving every leaf whose predicted norm exceeds eps its children
// (example/02_hyperbolic_05_burgers_adapt.cpp:246-291, DGAdapt::refine,
// source/DGAdapt.cpp:371-404).  The proposer replaces that prediction by a
// small network on surplus norms of the current solution; the threshold test
// stays as judge: after the RK step, leaves above eps that still lack
// children are refined and their instances redo the step from u^n.
//
// Training kernels are FP32, thread per sample for the activations and thread
// per parameter for the weight gradients (sums over the batch in sample
// order: deterministic, no atomics).  Inference keeps weights and activations
// in FP16 and accumulates in FP32.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "context.hpp"

namespace adg {

namespace {

constexpr int kPB = 256;
inline unsigned pblocks(int64_t n) { return static_cast<unsigned>((n + kPB - 1) / kPB); }

constexpr int F = kPropIn, H = kPropHidden;
// parameter offsets
constexpr int oW1 = 0, oB1 = oW1 + H * F, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H, oB3 = oW3 + H;
static_assert(oB3 + 1 == kPropParams, "parameter layout");

template <class T>
void to_dev(DevBuf<T>& b, const T* h, size_t n, hipStream_t s) {
  b.ensure(std::max<size_t>(n, 1));
  if (n) hip_check(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s), "proposer upload");
}

// ---------------------------------------------------------------------------
// features

__global__ __launch_bounds__(kPB) void k_elem_norm(const double* u, float* norm, int64_t E, int K) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double s = 0.;
  for (int i = 0; i < K; ++i) s += u[e * K + i] * u[e * K + i];
  norm[e] = static_cast<float>(sqrt(s));
}

struct FeatArgs {
  const int32_t* slot[kMaxDim];
  const int32_t* fbase[kMaxDim];
  const int32_t* fmap[kMaxDim];
  const float* norm;
  const double* eps;    // per instance
  const int32_t* inst;  // instance of each element
  float* feats;
  int32_t* leaf;
  int64_t E;
  int D, nmax;
};

__device__ __forceinline__ float feat_of(float norm, double eps) {
  const float v = log10f(fmaxf(norm, 1e-30f) / static_cast<float>(eps)) * 0.25f;
  return fminf(fmaxf(v, -4.f), 4.f);
}

__global__ __launch_bounds__(kPB) void k_features(FeatArgs a) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  const double eps = a.eps[a.inst ? a.inst[e] : 0];
  float f[F];
#pragma unroll
  for (int i = 0; i < F; ++i) f[i] = 0.f;
  f[0] = feat_of(a.norm[e], eps);
  int have = 0, total = 0, lsum = 0, lmax = 0;
  for (int d = 0; d < a.D; ++d) {
    const int s = a.slot[d][e];
    const int l = s == 0 ? 0 : 32 - __clz(s);
    lsum += l;
    lmax = max(lmax, l);
    const int32_t* fm = a.fmap[d] + a.fbase[d][e];
    float nl = a.norm[e], nr = a.norm[e];
    if (l >= 2) {
      const int lo = 1 << (l - 1), cnt = lo;
      const int sl = lo + (s - lo - 1 + cnt) % cnt, sr = lo + (s - lo + 1) % cnt;
      const int el = fm[sl], er = fm[sr];
      nl = el >= 0 ? a.norm[el] : 0.f;
      nr = er >= 0 ? a.norm[er] : 0.f;
    }
    f[1 + 2 * d] = feat_of(nl, eps);
    f[2 + 2 * d] = feat_of(nr, eps);
    if (l == 0) {
      total += 1;
      have += fm[1] >= 0;
    } else if (l < a.nmax) {
      total += 2;
      have += (fm[2 * s] >= 0) + (fm[2 * s + 1] >= 0);
    }
  }
  f[1 + 2 * a.D] = static_cast<float>(lsum) / static_cast<float>(a.D * a.nmax);
  f[2 + 2 * a.D] = static_cast<float>(lmax) / static_cast<float>(a.nmax);
#pragma unroll
  for (int i = 0; i < F; ++i) a.feats[e * F + i] = f[i];
  a.leaf[e] = have < total ? 1 : 0;
}

// ---------------------------------------------------------------------------
// training: forward and backward thread per sample, gradients thread per parameter

__device__ __forceinline__ float sigmoidf(float z) { return 1.f / (1.f + expf(-z)); }

// logistic loss with logits: -y log p - (1 - y) log(1 - p)
__device__ __forceinline__ float bce_logits(float z, float y) {
  const float sp = fmaxf(z, 0.f) + log1pf(expf(-fabsf(z)));  // log(1 + e^z)
  return sp - y * z;
}

__global__ __launch_bounds__(kPB) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ x,
                                                      const float* __restrict__ y, const float* __restrict__ w,
                                                      const int32_t* __restrict__ idx, int64_t n, float* h1,
                                                      float* h2, float* dz, float* loss) {
  __shared__ float sp[kPropParams];
  for (int i = threadIdx.x; i < kPropParams; i += blockDim.x) sp[i] = P[i];
  __syncthreads();
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int64_t src = idx ? idx[b] : b;
  float xi[F];
#pragma unroll
  for (int i = 0; i < F; ++i) xi[i] = x[src * F + i];
  float a1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float acc = sp[oB1 + j];
#pragma unroll
    for (int i = 0; i < F; ++i) acc = fmaf(sp[oW1 + j * F + i], xi[i], acc);
    a1[j] = fmaxf(acc, 0.f);
    h1[b * H + j] = a1[j];
  }
  float z = sp[oB3];
  for (int j = 0; j < H; ++j) {
    float acc = sp[oB2 + j];
#pragma unroll
    for (int k = 0; k < H; ++k) acc = fmaf(sp[oW2 + j * H + k], a1[k], acc);
    const float a2 = fmaxf(acc, 0.f);
    h2[b * H + j] = a2;
    z = fmaf(sp[oW3 + j], a2, z);
  }
  const float yy = y[src], ww = w ? w[src] : 1.f;
  dz[b] = ww * (sigmoidf(z) - yy);  // scaled by 1 / sum of weights on the host side of the gradient
  loss[b] = ww * bce_logits(z, yy);
}

__global__ __launch_bounds__(kPB) void k_mlp_backward(const float* __restrict__ P, const float* __restrict__ h1,
                                                       const float* __restrict__ h2, const float* __restrict__ dz,
                                                       int64_t n, float* dh1, float* dh2) {
  __shared__ float sp[kPropParams];
  for (int i = threadIdx.x; i < kPropParams; i += blockDim.x) sp[i] = P[i];
  __syncthreads();
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const float g = dz[b];
  float d2[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    d2[j] = h2[b * H + j] > 0.f ? g * sp[oW3 + j] : 0.f;
    dh2[b * H + j] = d2[j];
  }
  for (int k = 0; k < H; ++k) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < H; ++j) acc = fmaf(sp[oW2 + j * H + k], d2[j], acc);
    dh1[b * H + k] = h1[b * H + k] > 0.f ? acc : 0.f;
  }
}

// gradient of every parameter: sum over the batch in sample order, times scale
__global__ __launch_bounds__(kPB) void k_mlp_grad(const float* __restrict__ x, const int32_t* __restrict__ idx,
                                                   const float* __restrict__ h1, const float* __restrict__ h2,
                                                   const float* __restrict__ dz, const float* __restrict__ dh1,
                                                   const float* __restrict__ dh2, int64_t n, float scale,
                                                   float* grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kPropParams) return;
  float acc = 0.f;
  if (p < oB1) {
    const int j = p / F, i = p - j * F;
    for (int64_t b = 0; b < n; ++b) acc = fmaf(dh1[b * H + j], x[(idx ? idx[b] : b) * F + i], acc);
  } else if (p < oW2) {
    const int j = p - oB1;
    for (int64_t b = 0; b < n; ++b) acc += dh1[b * H + j];
  } else if (p < oB2) {
    const int j = (p - oW2) / H, k = (p - oW2) - j * H;
    for (int64_t b = 0; b < n; ++b) acc = fmaf(dh2[b * H + j], h1[b * H + k], acc);
  } else if (p < oW3) {
    const int j = p - oB2;
    for (int64_t b = 0; b < n; ++b) acc += dh2[b * H + j];
  } else if (p < oB3) {
    const int j = p - oW3;
    for (int64_t b = 0; b < n; ++b) acc = fmaf(dz[b], h2[b * H + j], acc);
  } else {
    for (int64_t b = 0; b < n; ++b) acc += dz[b];
  }
  grad[p] = acc * scale;
}

__global__ __launch_bounds__(kPB) void k_adam(float* P, const float* g, float* m1, float* m2, int n, float lr,
                                               float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float b1 = 0.9f, b2 = 0.999f;
  m1[p] = b1 * m1[p] + (1.f - b1) * g[p];
  m2[p] = b2 * m2[p] + (1.f - b2) * g[p] * g[p];
  P[p] -= lr * (m1[p] / c1) / (sqrtf(m2[p] / c2) + 1e-8f);
}

// ---------------------------------------------------------------------------
// inference: FP16 weights and activations, FP32 accumulation

__global__ __launch_bounds__(kPB) void k_to_half(const float* P, __half* Q, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) Q[p] = __float2half(P[p]);
}

__global__ __launch_bounds__(kPB) void k_mlp_infer(const __half* __restrict__ Q, const float* __restrict__ x,
                                                    int64_t n, float* prob) {
  __shared__ __half sq[kPropParams];
  for (int i = threadIdx.x; i < kPropParams; i += blockDim.x) sq[i] = Q[i];
  __syncthreads();
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= n) return;
  __half xi[F];
#pragma unroll
  for (int i = 0; i < F; ++i) xi[i] = __float2half(x[b * F + i]);
  __half a1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float acc = __half2float(sq[oB1 + j]);
#pragma unroll
    for (int i = 0; i < F; ++i) acc = fmaf(__half2float(sq[oW1 + j * F + i]), __half2float(xi[i]), acc);
    a1[j] = __float2half(fmaxf(acc, 0.f));
  }
  float z = __half2float(sq[oB3]);
  for (int j = 0; j < H; ++j) {
    float acc = __half2float(sq[oB2 + j]);
#pragma unroll
    for (int k = 0; k < H; ++k) acc = fmaf(__half2float(sq[oW2 + j * H + k]), __half2float(a1[k]), acc);
    const __half a2 = __float2half(fmaxf(acc, 0.f));
    z = fmaf(__half2float(sq[oW3 + j]), __half2float(a2), z);
  }
  prob[b] = sigmoidf(z);
}

// u = pred on the elements of instances whose mask is set
__global__ __launch_bounds__(kPB) void k_copy_masked(double* u, const double* pred, const int32_t* inst,
                                                      const int32_t* mask, int64_t n, int K) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mask[inst ? inst[i / K] : 0]) u[i] = pred[i];
}

// deterministic host generator (SplitMix64)
struct Mix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

// ---------------------------------------------------------------------------

void Context::proposer_features(const std::vector<double>& eps, float* feats_dev, int32_t* leaf_dev) {
  if (static_cast<int>(eps.size()) != B()) throw std::runtime_error("proposer features: one eps per instance");
  if (D > 4 || 1 + 2 * D + 2 > kPropIn) throw std::runtime_error("proposer features: dimension too large");
  if (E() == 0) return;
  prop_norm.ensure(E());
  to_dev(prop_eps, eps.data(), eps.size(), stream);
  hipLaunchKernelGGL(k_elem_norm, dim3(pblocks(E())), dim3(kPB), 0, stream, u.p, prop_norm.p, E(), K);
  FeatArgs a{};
  for (int d = 0; d < D; ++d) {
    a.slot[d] = g_slot[d].p;
    a.fbase[d] = g_fbase[d].p;
    a.fmap[d] = g_fmap[d].p;
  }
  a.norm = prop_norm.p;
  a.eps = prop_eps.p;
  a.inst = B() > 1 ? g_inst.p : nullptr;
  a.feats = feats_dev;
  a.leaf = leaf_dev;
  a.E = E();
  a.D = D;
  a.nmax = nmax;
  hipLaunchKernelGGL(k_features, dim3(pblocks(E())), dim3(kPB), 0, stream, a);
}

void Context::proposer_init(uint64_t seed) {
  // He initialisation of the weights, zero biases
  std::vector<float> p(kPropParams, 0.f);
  Mix g{seed};
  auto normal = [&]() {
    const double u1 = std::max(g.uniform(), 1e-300), u2 = g.uniform();
    return std::sqrt(-2. * std::log(u1)) * std::cos(2. * PI * u2);
  };
  for (int i = 0; i < H * F; ++i) p[oW1 + i] = static_cast<float>(normal() * std::sqrt(2.0 / F));
  for (int i = 0; i < H * H; ++i) p[oW2 + i] = static_cast<float>(normal() * std::sqrt(2.0 / H));
  for (int i = 0; i < H; ++i) p[oW3 + i] = static_cast<float>(normal() * std::sqrt(1.0 / H));
  proposer_set_params(p.data());
}

void Context::proposer_set_params(const float* host) {
  to_dev(prop_params, host, kPropParams, stream);
  std::vector<float> z(kPropParams, 0.f);
  to_dev(prop_m1, z.data(), kPropParams, stream);
  to_dev(prop_m2, z.data(), kPropParams, stream);
  prop_adam_t = 0;
  prop_ready = true;
  prop16_valid = false;
  sync();
}

void Context::proposer_get_params(float* host) {
  if (!prop_ready) throw std::runtime_error("proposer: no parameters (initialise or set them first)");
  hip_check(hipMemcpyAsync(host, prop_params.p, kPropParams * sizeof(float), hipMemcpyDeviceToHost, stream),
            "proposer download");
  sync();
}

double Context::proposer_gradient(const float* feats, const uint8_t* labels, const float* weights, int64_t n,
                                  float* grad) {
  if (!prop_ready) throw std::runtime_error("proposer: no parameters (initialise or set them first)");
  if (n < 1) throw std::runtime_error("proposer: no samples");
  std::vector<float> y(n), w(n);
  double wsum = 0.;
  for (int64_t i = 0; i < n; ++i) {
    y[i] = labels[i] ? 1.f : 0.f;
    w[i] = weights ? weights[i] : 1.f;
    wsum += w[i];
  }
  to_dev(prop_x, feats, static_cast<size_t>(n) * F, stream);
  to_dev(prop_y, y.data(), n, stream);
  to_dev(prop_w, w.data(), n, stream);
  for (DevBuf<float>* b : {&prop_h1, &prop_h2, &prop_dh1, &prop_dh2}) b->ensure(static_cast<size_t>(n) * H);
  prop_dz.ensure(n);
  prop_loss.ensure(n);
  prop_grad.ensure(kPropParams);
  hipLaunchKernelGGL(k_mlp_forward, dim3(pblocks(n)), dim3(kPB), 0, stream, prop_params.p, prop_x.p, prop_y.p,
                     prop_w.p, nullptr, n, prop_h1.p, prop_h2.p, prop_dz.p, prop_loss.p);
  hipLaunchKernelGGL(k_mlp_backward, dim3(pblocks(n)), dim3(kPB), 0, stream, prop_params.p, prop_h1.p, prop_h2.p,
                     prop_dz.p, n, prop_dh1.p, prop_dh2.p);
  hipLaunchKernelGGL(k_mlp_grad, dim3(pblocks(kPropParams)), dim3(kPB), 0, stream, prop_x.p, nullptr, prop_h1.p,
                     prop_h2.p, prop_dz.p, prop_dh1.p, prop_dh2.p, n, static_cast<float>(1.0 / wsum), prop_grad.p);
  std::vector<float> loss(n);
  hip_check(hipMemcpyAsync(grad, prop_grad.p, kPropParams * sizeof(float), hipMemcpyDeviceToHost, stream),
            "proposer download");
  hip_check(hipMemcpyAsync(loss.data(), prop_loss.p, n * sizeof(float), hipMemcpyDeviceToHost, stream),
            "proposer download");
  sync();
  double l = 0.;
  for (float v : loss) l += v;
  return l / wsum;
}

double Context::proposer_train(const float* feats, const uint8_t* labels, int64_t n, int epochs, int batch, float lr,
                               uint64_t seed) {
  if (!prop_ready) proposer_init(seed ^ 0x5A17u);
  if (n < 1 || epochs < 1 || batch < 1) throw std::runtime_error("proposer: bad training sizes");
  // class-balanced weights: each class carries half of the loss
  int64_t npos = 0;
  for (int64_t i = 0; i < n; ++i) npos += labels[i] != 0;
  const double wp = npos ? 0.5 * n / npos : 0., wn = npos < n ? 0.5 * n / (n - npos) : 0.;
  std::vector<float> y(n), w(n);
  for (int64_t i = 0; i < n; ++i) {
    y[i] = labels[i] ? 1.f : 0.f;
    w[i] = static_cast<float>(labels[i] ? wp : wn);
  }
  to_dev(prop_x, feats, static_cast<size_t>(n) * F, stream);
  to_dev(prop_y, y.data(), n, stream);
  to_dev(prop_w, w.data(), n, stream);
  const int64_t bs = std::min<int64_t>(batch, n);
  for (DevBuf<float>* b : {&prop_h1, &prop_h2, &prop_dh1, &prop_dh2}) b->ensure(static_cast<size_t>(bs) * H);
  prop_dz.ensure(bs);
  prop_loss.ensure(bs);
  prop_grad.ensure(kPropParams);
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = static_cast<int32_t>(i);
  Mix g{seed};
  std::vector<float> loss(bs);
  double last = 0.;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[g.next() % (i + 1)]);
    to_dev(prop_idx, perm.data(), n, stream);
    double lsum = 0., wsum_ep = 0.;
    for (int64_t b0 = 0; b0 + bs <= n; b0 += bs) {
      const int32_t* idx = prop_idx.p + b0;
      double wsum = 0.;
      for (int64_t i = 0; i < bs; ++i) wsum += w[perm[b0 + i]];
      hipLaunchKernelGGL(k_mlp_forward, dim3(pblocks(bs)), dim3(kPB), 0, stream, prop_params.p, prop_x.p, prop_y.p,
                         prop_w.p, idx, bs, prop_h1.p, prop_h2.p, prop_dz.p, prop_loss.p);
      hipLaunchKernelGGL(k_mlp_backward, dim3(pblocks(bs)), dim3(kPB), 0, stream, prop_params.p, prop_h1.p,
                         prop_h2.p, prop_dz.p, bs, prop_dh1.p, prop_dh2.p);
      hipLaunchKernelGGL(k_mlp_grad, dim3(pblocks(kPropParams)), dim3(kPB), 0, stream, prop_x.p, idx, prop_h1.p,
                         prop_h2.p, prop_dz.p, prop_dh1.p, prop_dh2.p, bs, static_cast<float>(1.0 / wsum),
                         prop_grad.p);
      ++prop_adam_t;
      const float c1 = 1.f - std::pow(0.9f, static_cast<float>(prop_adam_t));
      const float c2 = 1.f - std::pow(0.999f, static_cast<float>(prop_adam_t));
      hipLaunchKernelGGL(k_adam, dim3(pblocks(kPropParams)), dim3(kPB), 0, stream, prop_params.p, prop_grad.p,
                         prop_m1.p, prop_m2.p, kPropParams, lr, c1, c2);
      if (ep == epochs - 1) {
        hip_check(hipMemcpyAsync(loss.data(), prop_loss.p, bs * sizeof(float), hipMemcpyDeviceToHost, stream),
                  "proposer download");
        sync();
        for (float v : loss) lsum += v;
        wsum_ep += wsum;
      }
    }
    if (ep == epochs - 1) last = wsum_ep > 0 ? lsum / wsum_ep : 0.;
  }
  prop16_valid = false;
  sync();
  return last;
}

void Context::proposer_predict(const float* feats_dev, int64_t n, float* prob_dev) {
  if (!prop_ready) throw std::runtime_error("proposer: no parameters (initialise, set or train them first)");
  if (!prop16_valid) {
    prop_params16.ensure(kPropParams);
    hipLaunchKernelGGL(k_to_half, dim3(pblocks(kPropParams)), dim3(kPB), 0, stream, prop_params.p,
                       reinterpret_cast<__half*>(prop_params16.p), kPropParams);
    prop16_valid = true;
  }
  if (n > 0)
    hipLaunchKernelGGL(k_mlp_infer, dim3(pblocks(n)), dim3(kPB), 0, stream,
                       reinterpret_cast<const __half*>(prop_params16.p), feats_dev, n, prob_dev);
}

namespace {
// leaf: fewer children present than possible (the test of DGAdapt::refine)
bool host_leaf(const HostGrid& g, int64_t e) {
  int32_t s[kMaxDim];
  int have = 0;
  for (int d = 0; d < g.D; ++d) {
    for (int q = 0; q < g.D; ++q) s[q] = g.slot[e * g.D + q];
    const int sd = s[d], l = level_of_slot(sd);
    if (l == 0) { s[d] = 1; have += g.find(s) >= 0; }
    else if (l < g.nmax) {
      s[d] = 2 * sd; have += g.find(s) >= 0;
      s[d] = 2 * sd + 1; have += g.find(s) >= 0;
    }
  }
  return have < g.num_total_children(e);
}
}  // namespace

int64_t Context::proposer_step(const std::function<void(int, int, double)>& stage, const std::vector<double>& eps,
                               double* dt_inst) {
  // with a deferring judge and refine/coarsen on the device, the whole step
  // runs there (same proposals, step and judge)
  if (judge_defer && device_adapt_fits(E())) {
    copy_to_predict();
    return proposer_step_device(stage, eps);
  }
  ensure_host_grid();
  const int nb = B();
  copy_to_predict();  // u^n
  // The judge holds the step to the rule of the prediction it replaces: every
  // leaf of the grid at the start of the step whose norm exceeds eps gets its
  // children (elements added during the step are judged at the next one).
  std::vector<FlatIndex> start_leaves(nb);
  parallel_for(nb, [&](int i) {
    const HostGrid& g = grids[i];
    start_leaves[i].reset(static_cast<size_t>(g.E) * 2 + 2);
    for (int64_t e = 0; e < g.E; ++e)
      if (host_leaf(g, e)) start_leaves[i].insert(HostGrid::pack(&g.slot[e * D], D), 0);
  });
  // proposals
  std::vector<double> score(E(), 0.);
  {
    HostTimer timer(*this, HOST_ADAPT);
    if (refine_mode == 1) {
      prop_feats.ensure(static_cast<size_t>(std::max<int64_t>(E(), 1)) * F);
      prop_leaf.ensure(std::max<int64_t>(E(), 1));
      prop_prob.ensure(std::max<int64_t>(E(), 1));
      proposer_features(eps, prop_feats.p, prop_leaf.p);
      proposer_predict(prop_feats.p, E(), prop_prob.p);
      std::vector<float> prob(E());
      if (E() > 0)
        hip_check(hipMemcpyAsync(prob.data(), prop_prob.p, E() * sizeof(float), hipMemcpyDeviceToHost, stream),
                  "proposer download");
      sync();
      for (int64_t e = 0; e < E(); ++e) score[e] = prob[e] > prop_threshold ? 1. : 0.;
    } else {
      std::vector<double> s(E());
      indicator_sumsq(s.data());
      for (int i = 0; i < nb; ++i)
        for (int64_t e = off[i]; e < off[i + 1]; ++e) score[e] = threshold_norm(s[e], eps[i]) > eps[i] ? 1. : 0.;
    }
    std::vector<std::vector<int32_t>> ns(nb);
    std::vector<int64_t> added(nb, 0);
    parallel_for(nb, [&](int i) {
      const std::vector<double> si(score.begin() + off[i], score.begin() + off[i + 1]);
      ns[i] = refine_slots(grids[i], si, 0.5, added[i]);
    });
    std::vector<char> changed(nb, 0);
    int64_t tot = 0;
    for (int i = 0; i < nb; ++i) { tot += added[i]; changed[i] = added[i] > 0; }
    stat_proposed += tot;
    if (tot > 0) set_grids(ns, true, &changed);
  }
  // step, judge, redo the instances the judge refined
  const std::vector<double> dt0(dt_inst, dt_inst + nb);
  std::vector<int32_t> redo(nb, 1);
  int64_t redone = 0;
  constexpr int kMaxPasses = 6;
  for (int pass = 0;; ++pass) {
    for (int i = 0; i < nb; ++i) dt_inst[i] = redo[i] ? dt0[i] : 0.;
    if (nb > 1) upload_inst_dt();
    rk_init();
    for (int s = 0; s < 3; ++s) stage(3, s, dt0[0]);
    if (pass + 1 == kMaxPasses) break;
    HostTimer timer(*this, HOST_ADAPT);
    std::vector<double> s(E());
    indicator_sumsq(s.data());
    std::vector<std::vector<int32_t>> ns(nb);
    std::vector<int64_t> added(nb, 0);
    parallel_for(nb, [&](int i) {
      if (!redo[i]) { ns[i] = grids[i].slot; return; }
      const HostGrid& g = grids[i];
      std::vector<double> si(g.E, 0.);
      for (int64_t e = 0; e < g.E; ++e)
        si[e] = (threshold_norm(s[off[i] + e], eps[i]) > eps[i] &&
                 start_leaves[i].find(HostGrid::pack(&g.slot[e * D], D)) >= 0) ? 1. : 0.;
      ns[i] = refine_slots(g, si, 0.5, added[i]);
    });
    std::vector<char> changed(nb, 0);
    int64_t tot = 0;
    for (int i = 0; i < nb; ++i) {
      changed[i] = added[i] > 0;
      redo[i] = changed[i];
      tot += added[i];
      if (!judge_defer) redone += changed[i];
    }
    if (tot == 0) break;
    if (judge_defer) {
      // the step stands; the missed leaves get their children now (with zero
      // coefficients, as refinement gives them) and the next step evolves them
      set_grids(ns, true, &changed);
      stat_deferred += tot;
      break;
    }
    set_grids(ns, true, &changed);
    // the instances the judge refined start again from u^n
    to_dev(inst_mask, redo.data(), nb, stream);
    hipLaunchKernelGGL(k_copy_masked, dim3(pblocks(E() * K)), dim3(kPB), 0, stream, u.p, pred.p,
                       nb > 1 ? g_inst.p : nullptr, inst_mask.p, E() * K, K);
  }
  for (int i = 0; i < nb; ++i) dt_inst[i] = dt0[i];
  if (nb > 1) upload_inst_dt();
  stat_redo += redone;
  return redone;
}

}  // namespace adg
