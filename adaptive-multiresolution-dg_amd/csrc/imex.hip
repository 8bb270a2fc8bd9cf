// Artificial viscosity as an assembled operator (viscosity options 1 and 2 of
// example/02_hyperbolic_05_burgers_adapt.cpp:343-420) and the IMEX43 time
// stepper that treats it implicitly.
//
// ArtificialViscosity::assemble_matrix (source/BilinearForm.cpp:1190-1349)
// sums, over the viscosity elements w, the matrix of
// coef * int_{supp w} grad u . grad v restricted to the elements related to w.
// Here the host builds the block pattern once per assembly (the related sets,
// one (row element, column element) block per coupled pair, and for each block
// the viscosity elements that contribute, in viscosity-element order) and the
// device computes every block entry from a small table of 1D integrals: one
// thread per entry, contributions summed in the reference's order, so the
// values equal the oracle's bit for bit.  The operator is kept in block-sparse
// form (K x K blocks, K = (k+1)^D, rows by element).  rhs += (A u) is a
// thread-per-row product whose rows accumulate in increasing column order, the
// order of Eigen's column-major sparse product (bit-exact); the products inside
// the implicit solves, whose result is fixed only to the solver's tolerance,
// give each row a workgroup that splits the row's entries over its threads and
// adds the partial sums in a fixed tree (deterministic, and with hundreds of
// entries per row far more parallel than a thread per row).
//
// IMEX43 (source/ODESolver.cpp:872-1070) solves (I - dt alpha A) x = b four
// times per step with Eigen's ConjugateGradient (Jacobi preconditioner,
// tolerance 1e-12, at most 1000 iterations, zero initial guess).  The solve runs
// on the device for every instance at once: per-instance dot products
// (deterministic two-level sums over fixed chunks), per-instance scalars and
// convergence flags, read back every fourth iteration.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>

#include "context.hpp"

namespace adg {

namespace {

constexpr int kVB = 256;
inline unsigned vblocks(int64_t n) { return static_cast<unsigned>((n + kVB - 1) / kVB); }
constexpr int kChunk = kVB * 16;  // DoFs per reduction chunk
constexpr double kImexAlpha = 0.24169426078821, kImexBeta = 0.06042356519705, kImexEta = 0.12915286960590;
constexpr double kCgTol = 1e-12;
constexpr int kCgMaxIters = 1000;

template <class T>
void to_dev(DevBuf<T>& b, const std::vector<T>& h, hipStream_t s) {
  b.ensure(std::max<size_t>(h.size(), 1));
  if (!h.empty()) hip_check(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

// One entry of the operator per thread: blocks x K x K.
__global__ __launch_bounds__(kVB) void k_vop_assemble(int64_t n, int K, int kp, int D, const int32_t* recptr,
                                                      const int32_t* rec, const double* tab, double coef,
                                                      double* val) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int KK = K * K, kp2 = kp * kp;
  const int64_t b = t / KK;
  const int ij = static_cast<int>(t - b * KK);
  int i = ij / K, j = ij - (ij / K) * K;
  int it[kMaxDim], js[kMaxDim];
  for (int d = D - 1; d >= 0; --d) {
    it[d] = i % kp;
    js[d] = j % kp;
    i /= kp;
    j /= kp;
  }
  double acc = 0.;
  for (int r = recptr[b]; r < recptr[b + 1]; ++r) {
    double mat = 0.;
    for (int dd = 0; dd < D; ++dd) {
      double md = coef;
      for (int d = 0; d < D; ++d) {
        const double* tb = tab + static_cast<size_t>(rec[static_cast<size_t>(r) * D + d]) * 2 * kp2;
        md *= (d == dd) ? tb[kp2 + it[d] * kp + js[d]] : tb[it[d] * kp + js[d]];
      }
      mat += md;
    }
    if (fabs(mat) >= ROUND_OFF) acc += mat;
  }
  val[t] = acc;
}

// y = A x (mode 0), y += (A x) (mode 1), y = (I - c A) x (mode 2).
__global__ __launch_bounds__(kVB) void k_vop_spmv(int64_t n, int K, const int32_t* rowptr, const int32_t* col,
                                                  const double* val, const double* x, double* y, int mode,
                                                  double c) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t e = r / K;
  const int i = static_cast<int>(r - e * K);
  const int b0 = rowptr[e], b1 = rowptr[e + 1];
  double t = 0.;
  for (int b = b0; b < b1; ++b) {
    const int64_t ce = col[b];
    const double* v = val + (static_cast<size_t>(b) * K + i) * K;
    const double* xv = x + ce * K;
    if (mode == 2) {
      for (int j = 0; j < K; ++j) {
        const double m = (ce == e && j == i) ? 1. - c * v[j] : 0. - c * v[j];
        t += m * xv[j];
      }
    } else {
      for (int j = 0; j < K; ++j) t += v[j] * xv[j];
    }
  }
  if (mode == 0) y[r] = t;
  else if (mode == 1) y[r] = y[r] + t;
  else y[r] = (b0 == b1) ? x[r] : t;
}

// y = A x (mode 0) or y = (I - c A) x (mode 2): a workgroup per row, the
// row's entries split over the threads, partial sums added in a fixed tree.
constexpr int kWB = 256;
__global__ __launch_bounds__(kWB) void k_vop_spmv_wide(int K, const int32_t* rowptr, const int32_t* col,
                                                       const double* val, const double* x, double* y, int mode,
                                                       double c) {
  __shared__ double red[kWB];
  const int64_t r = blockIdx.x;
  const int64_t e = r / K;
  const int i = static_cast<int>(r - e * K);
  const int b0 = rowptr[e], b1 = rowptr[e + 1];
  const int nent = (b1 - b0) * K;
  double t = 0.;
  for (int q = threadIdx.x; q < nent; q += kWB) {
    const int bq = q / K, j = q - bq * K;
    const int b = b0 + bq;
    const int64_t ce = col[b];
    const double a = val[(static_cast<size_t>(b) * K + i) * K + j];
    const double m = (mode == 2) ? ((ce == e && j == i) ? 1. - c * a : 0. - c * a) : a;
    t += m * x[ce * K + j];
  }
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = kWB / 2; w > 0; w /= 2) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[r] = (mode == 2 && b0 == b1) ? x[r] : red[0];
}

// Jacobi preconditioner of I - c A: 1 / diagonal when non-zero, else 1.
__global__ __launch_bounds__(kVB) void k_vop_invdiag(int64_t n, int K, const int32_t* rowptr, const int32_t* col,
                                                     const double* val, double c, double* inv) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t e = r / K;
  const int i = static_cast<int>(r - e * K);
  double m = 1.;
  for (int b = rowptr[e]; b < rowptr[e + 1]; ++b)
    if (col[b] == e) m = 1. - c * val[(static_cast<size_t>(b) * K + i) * K + i];
  inv[r] = (m != 0.) ? 1. / m : 1.;
}

// Per-chunk partial sums (each chunk lies inside one instance): mode 0 sums
// a*b into slot 0; mode 1 sums a*a into slot 0 and a*(inv*a) into slot 1.
__global__ __launch_bounds__(kVB) void k_cg_dot(const int64_t* chunk, const double* a, const double* b,
                                                const double* inv, int mode, double* part) {
  __shared__ double s0[kVB], s1[kVB];
  const int64_t lo = chunk[2 * blockIdx.x], hi = chunk[2 * blockIdx.x + 1];
  double t0 = 0., t1 = 0.;
  for (int64_t q = lo + threadIdx.x; q < hi; q += kVB) {
    if (mode == 0) {
      t0 += a[q] * b[q];
    } else {
      t0 += a[q] * a[q];
      t1 += a[q] * (inv[q] * a[q]);
    }
  }
  s0[threadIdx.x] = t0;
  s1[threadIdx.x] = t1;
  __syncthreads();
  for (int w = kVB / 2; w > 0; w /= 2) {
    if (threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

// Per-instance CG scalars: [0] |b|^2, [1] threshold, [2] absNew, [3] |r|^2,
// [4] p.Mp, [5] beta, [6] dot slot 0, [7] dot slot 1; state [0] active,
// [1] iterations.
constexpr int kS = 8;

__global__ void k_cg_sum(int B, const int32_t* inst_chunk, const double* part, double* scal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  double a = 0., b = 0.;
  for (int c = inst_chunk[i]; c < inst_chunk[i + 1]; ++c) {
    a += part[2 * c];
    b += part[2 * c + 1];
  }
  scal[i * kS + 6] = a;
  scal[i * kS + 7] = b;
}

// after the first dot pass (r = b): |b|^2, threshold, |r|^2, absNew = r.(inv r)
__global__ void k_cg_begin(int B, double* scal, int32_t* state) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  double* s = scal + i * kS;
  const double rhs_norm2 = s[6];
  s[0] = rhs_norm2;
  s[1] = fmax(kCgTol * kCgTol * rhs_norm2, DBL_MIN);
  s[3] = s[6];
  s[2] = s[7];
  state[2 * i + 1] = 0;
  state[2 * i] = (rhs_norm2 == 0. || s[3] < s[1]) ? 0 : 1;
}

// x = 0, r = b, p = inv * r
__global__ __launch_bounds__(kVB) void k_cg_start(int64_t n, const double* b, const double* inv, double* x,
                                                  double* r, double* p) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  x[q] = 0.;
  r[q] = b[q];
  p[q] = inv[q] * b[q];
}

// x += alpha p, r -= alpha t with alpha = absNew / p.Mp of the row's instance
__global__ __launch_bounds__(kVB) void k_cg_update(int64_t n, int K, const int32_t* inst, const int32_t* state,
                                                   const double* scal, const double* p, const double* t, double* x,
                                                   double* r) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = inst[q / K];
  if (!state[2 * i]) return;
  const double alpha = scal[i * kS + 2] / scal[i * kS + 6];
  x[q] += alpha * p[q];
  r[q] -= alpha * t[q];
}

// after the residual dots: converged, or beta and the next absNew
__global__ void k_cg_step(int B, double* scal, int32_t* state) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B || !state[2 * i]) return;
  double* s = scal + i * kS;
  s[3] = s[6];
  if (s[3] < s[1]) {
    state[2 * i] = 0;
    return;
  }
  const double abs_old = s[2];
  s[2] = s[7];
  s[5] = s[2] / abs_old;
  if (++state[2 * i + 1] >= kCgMaxIters) state[2 * i] = 0;
}

// p = inv * r + beta p
__global__ __launch_bounds__(kVB) void k_cg_dir(int64_t n, int K, const int32_t* inst, const int32_t* state,
                                                const double* scal, const double* inv, const double* r, double* p) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = inst[q / K];
  if (!state[2 * i]) return;
  const double z = inv[q] * r[q];
  p[q] = z + scal[i * kS + 5] * p[q];
}

// right-hand sides of the implicit solves and the final combination
// (IMEX43::step_stage, source/ODESolver.cpp:994-1053, Eigen's evaluation order)
__global__ __launch_bounds__(kVB) void k_imex_combine(int64_t n, int stage, double dt, const double* tn,
                                                      const double* G0, const double* G1, const double* G2,
                                                      const double* G3, const double* F0, const double* F1,
                                                      const double* rhs, double* out) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double a = kImexAlpha;
  double v;
  if (stage == 0) {
    v = tn[q];
  } else if (stage == 1) {
    v = tn[q] + dt * (-a) * G0[q];
  } else if (stage == 2) {
    v = tn[q] + dt * rhs[q] + dt * (1 - a) * G1[q];
  } else if (stage == 3) {
    v = tn[q] + dt * 0.25 * F0[q] + dt * 0.25 * rhs[q] + dt * kImexBeta * G0[q] + dt * kImexEta * G1[q] +
        dt * (0.5 - kImexBeta - kImexEta - a) * G2[q];
  } else {
    v = tn[q] + dt * (1. / 6. * F0[q] + 1. / 6. * F1[q] + 2. / 3. * rhs[q]) +
        dt * (1. / 6. * G1[q] + 1. / 6. * G2[q] + 2. / 3. * G3[q]);
  }
  out[q] = v;
}

// 1D: do the supports of slots a and b intersect (open intervals; levels 0
// and 1 both span [0, 1])?
inline bool meet1(int32_t sa, int32_t sb) {
  const int la = level_of_slot(sa), lb = level_of_slot(sb);
  if (la <= 1 || lb <= 1) return true;
  const int32_t ca = sa - (1 << (la - 1)), cb = sb - (1 << (lb - 1));
  return la <= lb ? (cb >> (lb - la)) == ca : (ca >> (la - lb)) == cb;
}

}  // namespace

// ArtificialViscosity::assemble_matrix(coef, option) for the viscosity
// elements of the last update, every instance.
int64_t Context::visc_operator_assemble(double coef, int option) {
  ensure_host_grid();
  if (option != 1 && option != 2) throw std::runtime_error("viscosity operator option must be 1 or 2");
  const int L = nmax + 1, kp = k + 1;
  const int64_t nE = E();
  // viscosity elements by instance, in element (key) order; related sets
  struct WItem { int inst; int32_t le; };
  std::vector<WItem> ws;
  for (int i = 0; i < B(); ++i) {
    if (i >= static_cast<int>(visc_keys.size())) break;
    for (uint64_t key : visc_keys[i]) {
      const int32_t le = grids[i].index.find(key);
      if (le >= 0) ws.push_back({i, le});
    }
  }
  const int nw = static_cast<int>(ws.size());
  std::vector<std::vector<int32_t>> rel(nw);
  parallel_for(nw, [&](int q) {
    const HostGrid& g = grids[ws[q].inst];
    const int32_t* sw = &g.slot[static_cast<size_t>(ws[q].le) * D];
    std::vector<int32_t>& s = rel[q];
    if (option == 1) {
      // ptr_vol_intp: supports intersect in every dimension
      for (int64_t e = 0; e < g.E; ++e) {
        bool ok = true;
        for (int d = 0; d < D && ok; ++d) ok = meet1(g.slot[e * D + d], sw[d]);
        if (ok) s.push_back(static_cast<int32_t>(e));
      }
    } else {
      // find_all_parents: w and every ancestor reached through present parents
      std::vector<int32_t> cur{ws[q].le};
      s.push_back(ws[q].le);
      while (!cur.empty()) {
        std::vector<int32_t> add;
        for (int32_t e : cur) {
          int32_t sl[kMaxDim];
          for (int d = 0; d < D; ++d) {
            for (int dd = 0; dd < D; ++dd) sl[dd] = g.slot[static_cast<size_t>(e) * D + dd];
            if (sl[d] == 0) continue;
            sl[d] /= 2;
            const int32_t p = g.find(sl);
            if (p >= 0) add.push_back(p);
          }
        }
        std::sort(add.begin(), add.end());
        add.erase(std::unique(add.begin(), add.end()), add.end());
        s.insert(s.end(), add.begin(), add.end());
        cur.swap(add);
      }
      std::sort(s.begin(), s.end());
      s.erase(std::unique(s.begin(), s.end()), s.end());
    }
  });
  // records (row element, column element, 1D integral keys), per viscosity element
  struct Rec { int32_t a, b; int32_t key[kMaxDim]; };
  std::vector<std::vector<Rec>> recs(nw);
  parallel_for(nw, [&](int q) {
    const HostGrid& g = grids[ws[q].inst];
    const int64_t o = off[ws[q].inst];
    const int32_t* sw = &g.slot[static_cast<size_t>(ws[q].le) * D];
    const std::vector<int32_t>& s = rel[q];
    std::vector<Rec>& out = recs[q];
    for (int32_t a : s) {
      for (int32_t b : s) {
        Rec r{static_cast<int32_t>(o + a), static_cast<int32_t>(o + b), {0, 0, 0, 0}};
        bool ok = true;
        for (int d = 0; d < D && ok; ++d) {
          const int32_t sa = g.slot[static_cast<size_t>(a) * D + d], sb = g.slot[static_cast<size_t>(b) * D + d];
          ok = meet1(sa, sb);  // a and b meet w by construction of the related sets
          if (!ok) break;
          const int la = level_of_slot(sa), lb = level_of_slot(sb), lw = level_of_slot(sw[d]);
          int32_t f = sa;
          if (lb > level_of_slot(f)) f = sb;
          if (lw > level_of_slot(f)) f = sw[d];
          r.key[d] = ((f * L + la) * L + lb) * L + lw;
        }
        if (ok) out.push_back(r);
      }
    }
  });
  // 1D integral table: one entry (mass and stiffness, (k+1)^2 each) per key in use
  const int64_t nkeys = static_cast<int64_t>(1 << nmax) * L * L * L;
  if (static_cast<int64_t>(vop_key_tid.size()) != nkeys) {
    vop_key_tid.assign(nkeys, -1);
    vop_tab_host.clear();
  }
  const size_t tab_before = vop_tab_host.size();
  int64_t nrec = 0;
  for (auto& v : recs) {
    nrec += static_cast<int64_t>(v.size());
    for (Rec& r : v) {
      for (int d = 0; d < D; ++d) {
        int32_t& tid = vop_key_tid[r.key[d]];
        if (tid < 0) {
          // decode the key back into the three slots: ancestors of the finest slot
          int32_t key = r.key[d];
          const int lw = key % L; key /= L;
          const int lb = key % L; key /= L;
          const int la = key % L; key /= L;
          const int32_t f = key;
          const int lf = level_of_slot(f);
          auto anc = [&](int l) -> int32_t {
            if (l == 0) return 0;
            return (1 << (l - 1)) + ((f - (1 << (lf - 1))) >> (lf - l));
          };
          const int32_t sa = anc(la), sb = anc(lb), sw = anc(lw);
          tid = static_cast<int32_t>(vop_tab_host.size() / (2 * kp * kp));
          for (int pass = 0; pass < 2; ++pass)
            for (int pa = 0; pa < kp; ++pa)
              for (int pb = 0; pb < kp; ++pb)
                vop_tab_host.push_back(product_volume_interv(alpt, sa * kp + pa, sb * kp + pb, sw * kp, pass, pass,
                                                             pass == 0 ? k + 1 : k));
        }
        r.key[d] = tid;
      }
    }
  }
  if (nrec > INT32_MAX) throw std::runtime_error("viscosity operator: too many element pairs");
  // rows by element (counting sort, stable in viscosity-element order), then
  // columns within a row (stable): blocks with their records in order
  std::vector<int64_t> cnt(nE + 1, 0);
  for (auto& v : recs)
    for (const Rec& r : v) cnt[r.a + 1]++;
  for (int64_t e = 0; e < nE; ++e) cnt[e + 1] += cnt[e];
  std::vector<Rec> sorted(nrec);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (auto& v : recs)
      for (const Rec& r : v) sorted[pos[r.a]++] = r;
  }
  recs.clear();
  parallel_for(static_cast<int>(std::min<int64_t>(nE, 1 << 30)), [&](int e) {
    std::stable_sort(sorted.begin() + cnt[e], sorted.begin() + cnt[e + 1],
                     [](const Rec& x, const Rec& y) { return x.b < y.b; });
  });
  std::vector<int32_t> rowptr(nE + 1, 0), col, recptr{0}, rec;
  rec.reserve(static_cast<size_t>(nrec) * D);
  for (int64_t e = 0; e < nE; ++e) {
    for (int64_t q = cnt[e]; q < cnt[e + 1]; ++q) {
      if (q == cnt[e] || sorted[q].b != sorted[q - 1].b) {
        if (q != cnt[e]) recptr.push_back(static_cast<int32_t>(q));
        col.push_back(sorted[q].b);
      }
      for (int d = 0; d < D; ++d) rec.push_back(sorted[q].key[d]);
    }
    if (cnt[e + 1] > cnt[e]) recptr.push_back(static_cast<int32_t>(cnt[e + 1]));
    rowptr[e + 1] = static_cast<int32_t>(col.size());
  }
  vop_blocks = static_cast<int64_t>(col.size());
  to_dev(vop_rowptr, rowptr, stream);
  to_dev(vop_col, col, stream);
  to_dev(vop_recptr, recptr, stream);
  to_dev(vop_rec, rec, stream);
  if (vop_tab_host.size() != tab_before || vop_tab.cap < vop_tab_host.size()) to_dev(vop_tab, vop_tab_host, stream);
  const int64_t nval = vop_blocks * K * K;
  vop_val.ensure(std::max<int64_t>(nval, 1));
  if (nval > 0)
    hipLaunchKernelGGL(k_vop_assemble, dim3(vblocks(nval)), dim3(kVB), 0, stream, nval, K, kp, D, vop_recptr.p,
                       vop_rec.p, vop_tab.p, coef, vop_val.p);
  hip_check(hipGetLastError(), "viscosity operator assembly");
  hip_check(hipStreamSynchronize(stream), "viscosity operator assembly");
  vop_version = grid_version;
  vop_nnz_entries = nrec;
  return vop_blocks;
}

void Context::visc_operator_check() const {
  if (vop_version != grid_version)
    throw std::runtime_error("the viscosity operator was assembled on another grid (assemble it after refine/coarsen)");
}

void Context::visc_operator_apply(const double* x, double* y, int mode, double c) {
  const int64_t n = E() * K;
  if (n == 0) return;
  if (mode == 1)
    hipLaunchKernelGGL(k_vop_spmv, dim3(vblocks(n)), dim3(kVB), 0, stream, n, K, vop_rowptr.p, vop_col.p, vop_val.p,
                       x, y, mode, c);
  else
    hipLaunchKernelGGL(k_vop_spmv_wide, dim3(static_cast<unsigned>(n)), dim3(kWB), 0, stream, K, vop_rowptr.p,
                       vop_col.p, vop_val.p, x, y, mode, c);
}

void Context::rhs_visc_operator() {
  visc_operator_check();
  prof_begin(PROF_OTHER, 8.0 * (static_cast<double>(vop_blocks) * K * K + 3.0 * E() * K));
  visc_operator_apply(u.p, rhs.p, 1, 0.);
  prof_end();
}

void Context::imex43_init(double dt) {
  visc_operator_check();
  const int64_t n = E() * K;
  imex_dt = dt;
  imex_version = grid_version;
  for (DevBuf<double>* b : {&imex_tn, &imex_G[0], &imex_G[1], &imex_G[2], &imex_G[3], &imex_F[0], &imex_F[1],
                            &cg_b, &cg_r, &cg_p, &cg_t, &cg_inv})
    b->ensure(std::max<int64_t>(n, 1));
  if (n > 0) {
    hip_check(hipMemcpyAsync(imex_tn.p, u.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "IMEX43 init");
    hipLaunchKernelGGL(k_vop_invdiag, dim3(vblocks(n)), dim3(kVB), 0, stream, n, K, vop_rowptr.p, vop_col.p,
                       vop_val.p, dt * kImexAlpha, cg_inv.p);
  }
  // reduction chunks: fixed pieces of each instance's DoF range
  std::vector<int64_t> chunk;
  std::vector<int32_t> inst_chunk{0};
  for (int i = 0; i < B(); ++i) {
    for (int64_t lo = off[i] * K; lo < off[i + 1] * K; lo += kChunk) {
      chunk.push_back(lo);
      chunk.push_back(std::min<int64_t>(lo + kChunk, off[i + 1] * K));
    }
    inst_chunk.push_back(static_cast<int32_t>(chunk.size() / 2));
  }
  cg_nchunk = static_cast<int>(chunk.size() / 2);
  to_dev(cg_chunk, chunk, stream);
  to_dev(cg_inst_chunk, inst_chunk, stream);
  cg_part.ensure(std::max(2 * cg_nchunk, 1));
  cg_scal.ensure(static_cast<size_t>(B()) * kS);
  cg_state.ensure(static_cast<size_t>(B()) * 2);
  hip_check(hipStreamSynchronize(stream), "IMEX43 init");
}

// Conjugate gradient on (I - c A) x = cg_b for every instance; x is u.
// Returns the largest iteration count, -1 when an instance ends above the
// tolerance.
int Context::cg_solve(double c) {
  const int64_t n = E() * K;
  const int nB = B();
  if (n == 0) return 0;
  auto dots = [&](const double* a, const double* b, int mode) {
    if (cg_nchunk > 0)
      hipLaunchKernelGGL(k_cg_dot, dim3(cg_nchunk), dim3(kVB), 0, stream, cg_chunk.p, a, b, cg_inv.p, mode,
                         cg_part.p);
    hipLaunchKernelGGL(k_cg_sum, dim3((nB + 63) / 64), dim3(64), 0, stream, nB, cg_inst_chunk.p, cg_part.p,
                       cg_scal.p);
  };
  hipLaunchKernelGGL(k_cg_start, dim3(vblocks(n)), dim3(kVB), 0, stream, n, cg_b.p, cg_inv.p, u.p, cg_r.p, cg_p.p);
  dots(cg_r.p, nullptr, 1);
  hipLaunchKernelGGL(k_cg_begin, dim3((nB + 63) / 64), dim3(64), 0, stream, nB, cg_scal.p, cg_state.p);
  std::vector<int32_t> state(2 * nB);
  // The flags are read every kCheck iterations: an instance that converged
  // in between is skipped by every update kernel, so the extra launches
  // change nothing.
  constexpr int kCheck = 4;
  for (int iter = 0; iter <= kCgMaxIters + kCheck; ++iter) {
    if (iter % kCheck == 0) {
      hip_check(hipMemcpyAsync(state.data(), cg_state.p, state.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                               stream),
                "CG state");
      hip_check(hipStreamSynchronize(stream), "CG state");
      bool any = false;
      for (int i = 0; i < nB; ++i) any = any || state[2 * i];
      if (!any) break;
    }
    visc_operator_apply(cg_p.p, cg_t.p, 2, c);
    dots(cg_p.p, cg_t.p, 0);
    hipLaunchKernelGGL(k_cg_update, dim3(vblocks(n)), dim3(kVB), 0, stream, n, K, g_inst.p, cg_state.p, cg_scal.p,
                       cg_p.p, cg_t.p, u.p, cg_r.p);
    dots(cg_r.p, nullptr, 1);
    hipLaunchKernelGGL(k_cg_step, dim3((nB + 63) / 64), dim3(64), 0, stream, nB, cg_scal.p, cg_state.p);
    hipLaunchKernelGGL(k_cg_dir, dim3(vblocks(n)), dim3(kVB), 0, stream, n, K, g_inst.p, cg_state.p, cg_scal.p,
                       cg_inv.p, cg_r.p, cg_p.p);
  }
  hip_check(hipGetLastError(), "CG");
  std::vector<double> scal(static_cast<size_t>(nB) * kS);
  hip_check(hipMemcpyAsync(state.data(), cg_state.p, state.size() * sizeof(int32_t), hipMemcpyDeviceToHost, stream),
            "CG state");
  hip_check(hipMemcpyAsync(scal.data(), cg_scal.p, scal.size() * sizeof(double), hipMemcpyDeviceToHost, stream),
            "CG scalars");
  hip_check(hipStreamSynchronize(stream), "CG scalars");
  int worst = 0;
  bool fail = false;
  cg_iters.assign(nB, 0);
  for (int i = 0; i < nB; ++i) {
    const double* s = &scal[static_cast<size_t>(i) * kS];
    cg_iters[i] = state[2 * i + 1];
    worst = std::max(worst, state[2 * i + 1]);
    if (s[0] != 0. && !(std::sqrt(s[3] / s[0]) <= kCgTol)) fail = true;
  }
  return fail ? -1 : worst;
}

int Context::imex43_stage(int stage) {
  if (stage < 0 || stage > 4) throw std::runtime_error("IMEX43 stage must be 0..4");
  if (imex_version != grid_version || vop_version != grid_version)
    throw std::runtime_error("IMEX43: the grid changed since adg_imex43_init");
  const int64_t n = E() * K;
  if (n == 0) return 0;
  const double dt = imex_dt;
  double* out = (stage == 4) ? u.p : cg_b.p;
  hipLaunchKernelGGL(k_imex_combine, dim3(vblocks(n)), dim3(kVB), 0, stream, n, stage, dt, imex_tn.p, imex_G[0].p,
                     imex_G[1].p, imex_G[2].p, imex_G[3].p, imex_F[0].p, imex_F[1].p, rhs.p, out);
  if (stage == 2 || stage == 3)
    hip_check(hipMemcpyAsync(imex_F[stage - 2].p, rhs.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream),
              "IMEX43 explicit part");
  if (stage == 4) {
    hip_check(hipStreamSynchronize(stream), "IMEX43 final stage");
    return 0;
  }
  const int it = cg_solve(dt * kImexAlpha);
  visc_operator_apply(u.p, imex_G[stage].p, 0, 0.);
  hip_check(hipStreamSynchronize(stream), "IMEX43 stage");
  return it;
}

}  // namespace adg
