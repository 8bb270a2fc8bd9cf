// C-ABI entry points (include/adg.h) over the device context, the built-in
// point functions of the benchmark configurations, and the whole-run drivers.
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <unordered_map>
#include <memory>
#include <cstdio>

#include "../../include/adg.h"
#include "context.hpp"

struct adg_ctx {
  adg::Context* c = nullptr;
  std::string err;
};

namespace {

// Calls that read the host copy of the element tables get it rebuilt first
// (a device refine / coarsen leaves it stale); the ones marked kDeviceTables
// work from the device tables and leave it as it is.
enum GuardKind { kHostTables, kDeviceTables };

template <class F>
int guard(adg_ctx* h, F&& f, GuardKind kind = kHostTables) {
  if (!h || !h->c) return ADG_E_ARG;
  try {
    if (kind == kHostTables) h->c->ensure_host_grid();
    f(*h->c);
    return ADG_OK;
  } catch (const adg::HipError& e) {
    h->err = e.what();
    return ADG_E_HIP;
  } catch (const std::exception& e) {
    h->err = e.what();
    const std::string w = e.what();
    if (w.find("missing ancestor") != std::string::npos) return ADG_E_ANCESTOR;
    if (w.find("non-finite") != std::string::npos) return ADG_E_NONFINITE;
    return ADG_E_ARG;
  }
}

int field_size(const adg::Context& c, int field) {
  if (field == ADG_FIELD_UCOE || field == ADG_FIELD_RHS || field == ADG_FIELD_TN) return c.K;
  return c.P;
}

double* field_ptr(adg::Context& c, int field) {
  switch (field) {
    case ADG_FIELD_UCOE: return c.u.p;
    case ADG_FIELD_RHS: return c.rhs.p;
    case ADG_FIELD_UP: return c.up.p;
    case ADG_FIELD_TN: return c.tn.p;
    case ADG_FIELD_UINTP: return c.uintp.p;
  }
  if (field >= ADG_FIELD_FUCOE && field < ADG_FIELD_FUCOE + c.D)
    return c.fucoe.p + static_cast<int64_t>(c.same_flux ? 0 : field - ADG_FIELD_FUCOE) * c.E() * c.P;
  if (field >= ADG_FIELD_FP && field < ADG_FIELD_FP + c.D)
    return c.fp.p + static_cast<int64_t>(c.same_flux ? 0 : field - ADG_FIELD_FP) * c.E() * c.P;
  throw std::runtime_error("unknown field");
}

double periodic_dist2(const double* x, const double* c, int dim) {
  double r2 = 0.;
  for (int d = 0; d < dim; ++d) {
    double t = std::abs(x[d] - c[d]);
    t = std::min(t, 1. - t);
    r2 += t * t;
  }
  return r2;
}

adg::Context::StageProblem stage_problem(const adg_problem* p, int D) {
  adg::Context::StageProblem sp;
  sp.flux_kind = p->flux_kind;
  sp.flux_param = p->flux_param;
  for (int d = 0; d < D; ++d) sp.alpha[d] = p->alpha[d];
  return sp;
}

// Decision log of the predict-and-refine step: features of every element at
// the start of the step (kept for the leaves) ...
void log_features(adg::Context& c, const std::vector<double>& eps, std::vector<float>& lf, std::vector<int32_t>& ll) {
  const int64_t n = c.E();
  c.prop_feats.ensure(static_cast<size_t>(std::max<int64_t>(n, 1)) * adg::kPropIn);
  c.prop_leaf.ensure(std::max<int64_t>(n, 1));
  c.proposer_features(eps, c.prop_feats.p, c.prop_leaf.p);
  lf.resize(static_cast<size_t>(n) * adg::kPropIn);
  ll.resize(n);
  if (n > 0) {
    adg::hip_check(hipMemcpyAsync(lf.data(), c.prop_feats.p, lf.size() * sizeof(float), hipMemcpyDeviceToHost,
                                  c.stream), "download");
    adg::hip_check(hipMemcpyAsync(ll.data(), c.prop_leaf.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream),
                   "download");
  }
  c.sync();
}

// ... and whether the refinement that follows the prediction refines them
// (leaf with predicted norm above eps, DGAdapt::refine).
void log_labels(adg::Context& c, const std::vector<double>& eps, const std::vector<float>& lf,
                const std::vector<int32_t>& ll) {
  std::vector<double> s(c.E());
  c.indicator_sumsq(s.data());
  for (int i = 0; i < c.B(); ++i)
    for (int64_t e = c.off[i]; e < c.off[i + 1]; ++e) {
      if (!ll[e]) continue;
      c.log_feats.insert(c.log_feats.end(), &lf[e * adg::kPropIn], &lf[e * adg::kPropIn] + adg::kPropIn);
      c.log_labels.push_back(adg::threshold_norm(s[e], eps[i]) > eps[i] ? 1 : 0);
    }
}

// one RK3 stage of the Lagrange-interpolated hyperbolic operator
void hyperbolic_rhs(adg::Context& c, const adg_problem* p) {
  const double fp = p->flux_param;
  c.eval_up();
  c.flux(p->flux_kind, &fp);
  c.fp_to_coe();
  c.rhs_zero();
  c.rhs_lagr(false);
  c.rhs_lagr(true);
  c.rhs_penalty(p->alpha);
}

}  // namespace

namespace {
const char kStateMagic[8] = {'A', 'D', 'G', 'S', 'T', 'A', 'T', '1'};

struct FileCloser {
  void operator()(std::FILE* f) const { if (f) std::fclose(f); }
};
using FilePtr = std::unique_ptr<std::FILE, FileCloser>;

void write_all(std::FILE* f, const void* p, size_t n) {
  if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("state file: write failed");
}
void read_all(std::FILE* f, void* p, size_t n) {
  if (n && std::fread(p, 1, n, f) != n) throw std::runtime_error("state file: truncated");
}
}  // namespace

extern "C" {

int adg_create(int dim, int k, int m, int nmax, adg_ctx** out) {
  if (!out) return ADG_E_ARG;
  *out = nullptr;
  adg_ctx* h = new adg_ctx();
  try {
    h->c = new adg::Context(dim, k, m, nmax);
    adg::hip_check(hipGetDevice(&h->c->device), "hipGetDevice");
    adg::noalloc_setup(h->c->device);  // device address of the scratch array, once per GPU
  } catch (const adg::HipError& e) {
    h->err = e.what();
    *out = h;
    return ADG_E_HIP;
  } catch (const std::exception& e) {
    h->err = e.what();
    *out = h;
    return ADG_E_ARG;
  }
  *out = h;
  return ADG_OK;
}

int adg_destroy(adg_ctx* h) {
  if (!h) return ADG_E_ARG;
  delete h->c;
  delete h;
  return ADG_OK;
}

const char* adg_last_error(adg_ctx* h) { return h ? h->err.c_str() : "null context"; }

int adg_sync(adg_ctx* h) {
  return guard(h, [](adg::Context& c) {
    c.sync();
    c.check_error_flag();
  }, kDeviceTables);
}

int adg_grid_init(adg_ctx* h, int sparse, int level_init) {
  return guard(h, [&](adg::Context& c) {
    if (level_init < 0 || level_init > c.nmax) throw std::runtime_error("level_init must be in [0, nmax]");
    c.set_grid_slots(adg::HostGrid::initial_slots(c.D, level_init, sparse != 0), false);
    c.inst_eps.clear();
    c.inst_eta.clear();
    c.sync();
  });
}

int adg_grid_set(adg_ctx* h, const int32_t* level, const int32_t* suppt, int64_t n) {
  return guard(h, [&](adg::Context& c) {
    std::vector<int32_t> s(static_cast<size_t>(n) * c.D);
    for (int64_t i = 0; i < n * c.D; ++i) {
      if (level[i] < 0 || level[i] > c.nmax) throw std::runtime_error("level out of range");
      const int jmax = level[i] == 0 ? 1 : (1 << level[i]) - 1;
      if (suppt[i] < 1 || suppt[i] > jmax || suppt[i] % 2 == 0) throw std::runtime_error("support index out of range");
      s[i] = adg::slot_of(level[i], suppt[i]);
    }
    c.set_grid_slots(s, false);
    c.inst_eps.clear();
    c.inst_eta.clear();
    c.sync();
  });
}

int64_t adg_num_elem(adg_ctx* h) { return (h && h->c) ? h->c->E() : -1; }

int adg_grid_get(adg_ctx* h, int32_t* level, int32_t* suppt, int32_t* hash) {
  return guard(h, [&](adg::Context& c) {
    for (int64_t e = 0; e < c.E(); ++e) {
      for (int d = 0; d < c.D; ++d) {
        if (level) level[e * c.D + d] = c.level(e, d);
        if (suppt) suppt[e * c.D + d] = adg::suppt_of_slot(c.uslot[e * c.D + d]);
      }
      if (hash) hash[e] = c.uhash[e];
    }
  });
}

int adg_max_level(adg_ctx* h, int32_t* out) {
  return guard(h, [&](adg::Context& c) {
    for (int d = 0; d < c.D; ++d) out[d] = 0;
    for (int64_t e = 0; e < c.E(); ++e)
      for (int d = 0; d < c.D; ++d) out[d] = std::max<int32_t>(out[d], c.level(e, d));
  });
}

int32_t adg_hash_key(const int32_t* level, const int32_t* suppt, int dim) {
  static const adg::HashKey hk;
  return hk.key(level, suppt, dim);
}

int adg_field_set(adg_ctx* h, int field, const double* host) {
  return guard(h, [&](adg::Context& c) {
    if (field >= ADG_FIELD_FUCOE) c.same_flux = false;
    const size_t n = static_cast<size_t>(c.E()) * field_size(c, field);
    adg::hip_check(hipMemcpyAsync(field_ptr(c, field), host, n * sizeof(double), hipMemcpyHostToDevice, c.stream),
                   "field upload");
    c.sync();
  }, kDeviceTables);
}

int adg_field_get(adg_ctx* h, int field, double* host) {
  return guard(h, [&](adg::Context& c) {
    const size_t n = static_cast<size_t>(c.E()) * field_size(c, field);
    adg::hip_check(hipMemcpyAsync(host, field_ptr(c, field), n * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                   "field download");
    c.sync();
  }, kDeviceTables);
}

namespace {
// Argument checks and tensor sizes of a host-array sweep; returns (pf, pt).
void sweep_shape(const adg::Context& c, int table, int lu, int dim, const int32_t* size_in, const int32_t* size_out,
                 int* sin, int* sout, int64_t& nin, int64_t& nout, int& pf, int& pt) {
  if (dim < 0 || dim >= c.D) throw std::runtime_error("dimension out of range");
  if (lu < 0 || lu > 2) throw std::runtime_error("lu must be 0, 1 or 2");
  if (table < 0 || table > 8) throw std::runtime_error("unknown table id");
  pf = (table == 0 || table >= 4) ? c.k : c.m;
  pt = (table == 0) ? c.m : c.k;
  nin = 1;
  nout = 1;
  const int mx = std::max(c.k, c.m) + 1;
  for (int d = 0; d < c.D; ++d) {
    sin[d] = size_in[d];
    sout[d] = size_out[d];
    if (sin[d] < 1 || sin[d] > mx || sout[d] < 1 || sout[d] > mx) throw std::runtime_error("tensor size out of range");
    if (d != dim && sin[d] != sout[d]) throw std::runtime_error("sizes may differ only in the swept dimension");
    nin *= sin[d];
    nout *= sout[d];
  }
  if (sin[dim] != pf + 1 || sout[dim] != pt + 1) throw std::runtime_error("sweep: size mismatch in swept dimension");
}

// in -> out through device staging buffers din (E*nin) and dout (E*nout)
void sweep_staged(adg::Context& c, int table, int flx, int lu, int dim, const int* sin, const int* sout, int64_t nin,
                  int64_t nout, int pf, int pt, double coef, const double* in, double* out, double* din,
                  double* dout) {
  adg::hip_check(hipMemcpyAsync(din, in, c.E() * nin * sizeof(double), hipMemcpyHostToDevice, c.stream), "upload");
  c.sweep(din, sin, dout, nullptr, sout, dim, c.table(table), flx != 0, lu, coef, pf, pt);
  adg::hip_check(hipMemcpyAsync(out, dout, c.E() * nout * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                 "download");
  c.sync();
}
}  // namespace

int adg_sweep(adg_ctx* h, int table, int flx, int lu, int dim, const int32_t* size_in, const int32_t* size_out,
              double coef, const double* in, double* out) {
  return guard(h, [&](adg::Context& c) {
    int sin[adg::kMaxDim], sout[adg::kMaxDim], pf, pt;
    int64_t nin, nout;
    sweep_shape(c, table, lu, dim, size_in, size_out, sin, sout, nin, nout, pf, pt);
    c.tmpA.ensure(std::max<int64_t>(c.E() * nin, 1));
    c.tmpB.ensure(std::max<int64_t>(c.E() * nout, 1));
    sweep_staged(c, table, flx, lu, dim, sin, sout, nin, nout, pf, pt, coef, in, out, c.tmpA.p, c.tmpB.p);
  });
}

int adg_sweep_noalloc(adg_ctx* h, int table, int flx, int lu, int dim, const int32_t* size_in,
                      const int32_t* size_out, double coef, const double* in, double* out) {
  return guard(h, [&](adg::Context& c) {
    int sin[adg::kMaxDim], sout[adg::kMaxDim], pf, pt;
    int64_t nin, nout;
    sweep_shape(c, table, lu, dim, size_in, size_out, sin, sout, nin, nout, pf, pt);
    const int64_t need = c.E() * (nin + nout);
    if (need > static_cast<int64_t>(adg::kNoallocScratchDoubles))
      throw std::runtime_error("adg_sweep_noalloc: E x (input + output) = " + std::to_string(need) +
                               " values exceeds the scratch capacity of " +
                               std::to_string(adg::kNoallocScratchDoubles) + " (use adg_sweep)");
    double* s = adg::noalloc_scratch(c.device);  // the array's device address on this GPU
    sweep_staged(c, table, flx, lu, dim, sin, sout, nin, nout, pf, pt, coef, in, out, s, s + c.E() * nin);
  });
}

int adg_lagr_points(adg_ctx* h, double* x) {
  return guard(h, [&](adg::Context& c) {
    if (!x) throw std::runtime_error("null output");
    const int p1 = c.m + 1;
    for (int64_t e = 0; e < c.E(); ++e)
      for (int q = 0; q < c.P; ++q) {
        int rem = q;
        for (int d = c.D - 1; d >= 0; --d) {
          const int pd = rem % p1;
          rem /= p1;
          x[(e * c.P + q) * c.D + d] = c.interp_basis().point(c.uslot[e * c.D + d] * p1 + pd);
        }
      }
  });
}

int adg_viscosity_elements(adg_ctx* h, double shock_kappa, int64_t* count, int32_t* flags) {
  return guard(h, [&](adg::Context& c) {
    const int64_t n = c.viscosity_elements(shock_kappa);
    if (count) *count = n;
    if (flags) {
      std::fill(flags, flags + c.E(), 0);
      for (int64_t e : c.visc_elems) flags[e] = 1;
    }
  });
}

int adg_viscosity_flux(adg_ctx* h, double coefficient, int overlap) {
  return guard(h, [&](adg::Context& c) { c.viscosity_flux(coefficient, overlap != 0); });
}

int adg_viscosity_operator(adg_ctx* h, double coefficient, int option, int64_t* blocks) {
  return guard(h, [&](adg::Context& c) {
    const int64_t n = c.visc_operator_assemble(coefficient, option);
    if (blocks) *blocks = n;
  });
}

int adg_rhs_viscosity_operator(adg_ctx* h) {
  return guard(h, [&](adg::Context& c) { c.rhs_visc_operator(); });
}

int adg_imex43_init(adg_ctx* h, double dt) {
  return guard(h, [&](adg::Context& c) { c.imex43_init(dt); });
}

int adg_imex43_stage(adg_ctx* h, int stage, int* cg_iterations) {
  int it = 0;
  const int rc = guard(h, [&](adg::Context& c) { it = c.imex43_stage(stage); });
  if (rc != ADG_OK) return rc;
  if (cg_iterations) *cg_iterations = it;
  if (it < 0) {
    h->err = "IMEX43: conjugate gradient did not reach its tolerance";
    return ADG_E_NOCONVERGE;
  }
  return ADG_OK;
}

int adg_set_refine_mode(adg_ctx* h, int mode, double threshold) {
  return guard(h, [&](adg::Context& c) {
    if (mode != ADG_REFINE_PREDICT && mode != ADG_REFINE_LEARNED && mode != ADG_REFINE_CURRENT)
      throw std::runtime_error("unknown refine mode");
    if (!(threshold > 0. && threshold < 1.)) throw std::runtime_error("proposer threshold must lie in (0, 1)");
    c.refine_mode = mode;
    c.prop_threshold = threshold;
  });
}

int adg_set_judge_defer(adg_ctx* h, int defer) {
  return guard(h, [&](adg::Context& c) {
    if (defer != 0 && defer != 1) throw std::runtime_error("judge defer must be 0 or 1");
    c.judge_defer = defer;
  });
}

int adg_proposer_stats(adg_ctx* h, int64_t* proposed, int64_t* redone, int64_t* deferred) {
  return guard(h, [&](adg::Context& c) {
    if (proposed) *proposed = c.stat_proposed;
    if (redone) *redone = c.stat_redo;
    if (deferred) *deferred = c.stat_deferred;
  }, kDeviceTables);
}

int adg_set_device_adapt(adg_ctx* h, int on) {
  return guard(h, [&](adg::Context& c) {
    if (on != 0 && on != 1) throw std::runtime_error("device adapt must be 0 or 1");
    c.device_adapt = on != 0;
  }, kDeviceTables);
}

int adg_device_adapt_stats(adg_ctx* h, int64_t* device_calls, int64_t* host_fallbacks, int32_t* reasons) {
  return guard(h, [&](adg::Context& c) {
    if (device_calls) *device_calls = c.stat_dev_adapt;
    if (host_fallbacks) *host_fallbacks = c.stat_dev_fallback;
    if (reasons) *reasons = c.dev_fallback_flags;
  }, kDeviceTables);
}

int adg_grid_check(adg_ctx* h, int64_t* mismatches) {
  return guard(h, [&](adg::Context& c) {
    if (!mismatches) throw std::runtime_error("grid check: null output");
    *mismatches = c.check_device_grid();
  });
}

int adg_proposer_init(adg_ctx* h, uint64_t seed) { return guard(h, [&](adg::Context& c) { c.proposer_init(seed); }); }

int adg_proposer_set_params(adg_ctx* h, const float* params) {
  return guard(h, [&](adg::Context& c) {
    if (!params) throw std::runtime_error("null parameters");
    c.proposer_set_params(params);
  });
}

int adg_proposer_get_params(adg_ctx* h, float* params) {
  return guard(h, [&](adg::Context& c) {
    if (!params) throw std::runtime_error("null output");
    c.proposer_get_params(params);
  });
}

int adg_proposer_gradient(adg_ctx* h, const float* feats, const uint8_t* labels, const float* weights, int64_t n,
                          float* grad, double* loss) {
  return guard(h, [&](adg::Context& c) {
    if (!feats || !labels || !grad) throw std::runtime_error("null array");
    const double l = c.proposer_gradient(feats, labels, weights, n, grad);
    if (loss) *loss = l;
  });
}

int adg_proposer_train(adg_ctx* h, const float* feats, const uint8_t* labels, int64_t n, int epochs, int batch,
                       float lr, uint64_t seed, double* loss) {
  return guard(h, [&](adg::Context& c) {
    if (!feats || !labels) throw std::runtime_error("null array");
    const double l = c.proposer_train(feats, labels, n, epochs, batch, lr, seed);
    if (loss) *loss = l;
  });
}

int adg_proposer_predict(adg_ctx* h, const float* feats, int64_t n, float* prob) {
  return guard(h, [&](adg::Context& c) {
    if (!feats || !prob || n < 0) throw std::runtime_error("null array or negative count");
    c.prop_x.ensure(static_cast<size_t>(std::max<int64_t>(n, 1)) * adg::kPropIn);
    c.prop_prob.ensure(std::max<int64_t>(n, 1));
    if (n > 0)
      adg::hip_check(hipMemcpyAsync(c.prop_x.p, feats, n * adg::kPropIn * sizeof(float), hipMemcpyHostToDevice,
                                    c.stream), "upload");
    c.proposer_predict(c.prop_x.p, n, c.prop_prob.p);
    if (n > 0)
      adg::hip_check(hipMemcpyAsync(prob, c.prop_prob.p, n * sizeof(float), hipMemcpyDeviceToHost, c.stream),
                     "download");
    c.sync();
  });
}

int adg_proposer_features(adg_ctx* h, const double* eps, float* feats, int32_t* leaf) {
  return guard(h, [&](adg::Context& c) {
    if (!eps || !feats || !leaf) throw std::runtime_error("null array");
    std::vector<float> lf;
    std::vector<int32_t> ll;
    log_features(c, std::vector<double>(eps, eps + c.B()), lf, ll);
    std::copy(lf.begin(), lf.end(), feats);
    std::copy(ll.begin(), ll.end(), leaf);
  });
}

int adg_decision_log(adg_ctx* h, int enable) {
  return guard(h, [&](adg::Context& c) {
    c.log_decisions = enable != 0;
    if (enable) {
      c.log_feats.clear();
      c.log_labels.clear();
    }
  });
}

int64_t adg_decision_log_size(adg_ctx* h) {
  return (h && h->c) ? static_cast<int64_t>(h->c->log_labels.size()) : -1;
}

int adg_decision_log_get(adg_ctx* h, float* feats, uint8_t* labels) {
  return guard(h, [&](adg::Context& c) {
    if (feats) std::copy(c.log_feats.begin(), c.log_feats.end(), feats);
    if (labels) std::copy(c.log_labels.begin(), c.log_labels.end(), labels);
  });
}

int adg_decision_log_write(adg_ctx* h, const char* path) {
  return guard(h, [&](adg::Context& c) {
    if (!path) throw std::runtime_error("null path");
    FilePtr f(std::fopen(path, "wb"));
    if (!f) throw std::runtime_error(std::string("decision log: cannot open ") + path);
    const char magic[8] = {'A', 'D', 'G', 'D', 'L', 'O', 'G', '1'};
    const int32_t hdr[2] = {adg::kPropIn, 0};
    const int64_t n = static_cast<int64_t>(c.log_labels.size());
    write_all(f.get(), magic, sizeof(magic));
    write_all(f.get(), hdr, sizeof(hdr));
    write_all(f.get(), &n, sizeof(n));
    write_all(f.get(), c.log_feats.data(), c.log_feats.size() * sizeof(float));
    write_all(f.get(), c.log_labels.data(), c.log_labels.size());
    if (std::fflush(f.get()) != 0) throw std::runtime_error("decision log: write failed");
  });
}

int adg_set_precision(adg_ctx* h, int precision) {
  return guard(h, [&](adg::Context& c) {
    if (precision != ADG_PREC_FP64 && precision != ADG_PREC_FP32_SPLIT) throw std::runtime_error("unknown precision");
    c.precision = precision;
  });
}

int adg_set_interpolation(adg_ctx* h, int family) {
  return guard(h, [&](adg::Context& c) { c.set_interpolation(family); });
}

int adg_set_level_sum_cap(adg_ctx* h, int cap) {
  return guard(h, [&](adg::Context& c) { c.lsum_cap = cap < 0 ? -1 : cap; });
}
int adg_eval_up(adg_ctx* h) { return guard(h, [](adg::Context& c) { c.eval_up(); }, kDeviceTables); }
int adg_flux(adg_ctx* h, int kind, const double* params) {
  return guard(h, [&](adg::Context& c) {
    if (kind < 0 || kind > 2) throw std::runtime_error("unknown flux kind");
    c.flux(kind, params);
  });
}
int adg_flux_set(adg_ctx* h, int ncomp, const double* fp) {
  return guard(h, [&](adg::Context& c) {
    if (ncomp != 1 && ncomp != c.D) throw std::runtime_error("ncomp must be 1 or dim");
    const size_t n = static_cast<size_t>(c.E()) * c.P * ncomp;
    adg::hip_check(hipMemcpyAsync(c.fp.p, fp, n * sizeof(double), hipMemcpyHostToDevice, c.stream), "flux upload");
    c.same_flux = (ncomp == 1);
    c.sync();
  });
}
int adg_fp_to_coe(adg_ctx* h) {
  return guard(h, [](adg::Context& c) {
    c.fp_to_coe();
    c.check_error_flag();  // a missing ancestor (hole in the grid)
  }, kDeviceTables);
}
int adg_rhs_zero(adg_ctx* h) { return guard(h, [](adg::Context& c) { c.rhs_zero(); }, kDeviceTables); }
int adg_rhs_lagr_vol(adg_ctx* h) { return guard(h, [](adg::Context& c) { c.rhs_lagr(false); }, kDeviceTables); }
int adg_rhs_lagr_flx(adg_ctx* h) { return guard(h, [](adg::Context& c) { c.rhs_lagr(true); }, kDeviceTables); }
int adg_rhs_penalty(adg_ctx* h, const double* alpha) { return guard(h, [&](adg::Context& c) { c.rhs_penalty(alpha); }, kDeviceTables); }
int adg_rhs_linear(adg_ctx* h, const double* coef) { return guard(h, [&](adg::Context& c) { c.rhs_linear(coef); }); }
int adg_rk_init(adg_ctx* h) { return guard(h, [](adg::Context& c) { c.rk_init(); }, kDeviceTables); }
int adg_rk_stage(adg_ctx* h, int scheme, int stage, double dt) {
  return guard(h, [&](adg::Context& c) {
    if (!((scheme == ADG_RK_EULER && stage == 0) || (scheme == ADG_RK2 && stage >= 0 && stage <= 1) ||
          (scheme == ADG_RK3 && stage >= 0 && stage <= 2)))
      throw std::runtime_error("invalid RK scheme/stage");
    c.rk_stage(scheme, stage, dt);
  }, kDeviceTables);
}
int adg_copy_to_predict(adg_ctx* h) { return guard(h, [](adg::Context& c) { c.copy_to_predict(); }, kDeviceTables); }
int adg_copy_predict_to_ucoe(adg_ctx* h) { return guard(h, [](adg::Context& c) { c.copy_predict_to_ucoe(); }, kDeviceTables); }
int adg_indicator_sumsq(adg_ctx* h, double* out) { return guard(h, [&](adg::Context& c) { c.indicator_sumsq(out); }, kDeviceTables); }
int adg_refine(adg_ctx* h, double eps, int64_t* n_added) {
  return guard(h, [&](adg::Context& c) {
    const int64_t a = c.refine(eps);
    if (n_added) *n_added = a;
  }, kDeviceTables);
}
int adg_coarsen(adg_ctx* h, double eta, int64_t* n_removed) {
  return guard(h, [&](adg::Context& c) {
    const int64_t r = c.coarsen(eta);
    if (n_removed) *n_removed = r;
  }, kDeviceTables);
}
int adg_init_adaptive_intp(adg_ctx* h, adg_point_fn f, void* user, double eps) {
  return guard(h, [&](adg::Context& c) {
    if (!f) throw std::runtime_error("null point function");
    c.init_adaptive_intp(f, user, eps);
  });
}

double adg_ic_eval(const double* x, int dim, void* user) {
  const adg_ic_params* p = static_cast<const adg_ic_params*>(user);
  const double pi = adg::PI;
  switch (p->kind) {
    case 0: {
      double s = 0.;
      for (int d = 0; d < dim; ++d) s += x[d];
      return std::sin(2. * pi * (s - p->shift));
    }
    case 1: {
      double v = 1.;
      for (int d = 0; d < dim; ++d) v *= std::sin(2. * pi * (x[d] - p->shift));
      return v;
    }
    case 2:
    case 4: {
      double s = 0.;
      for (int d = 0; d < dim; ++d) s += x[d];
      double v = (p->kind == 2) ? p->base + p->ampl * std::sin(2. * pi * s) : 0.;
      for (int b = 0; b < p->nbump && b < 8; ++b) {
        const double r2 = periodic_dist2(x, p->bump_center[b], dim);
        v += p->bump_amp[b] * std::exp(-r2 / (p->bump_width[b] * p->bump_width[b]));
      }
      return v;
    }
    case 3: {
      double r2 = 0.;
      for (int d = 0; d < dim; ++d) {
        const double X = p->offset[d] + p->scale[d] * x[d] - p->center[d];
        r2 += X * X;
      }
      return (r2 <= p->radius * p->radius) ? p->inside : p->outside;
    }
  }
  return 0.;
}

int adg_init_from_points(adg_ctx* h, const double* values) {
  return guard(h, [&](adg::Context& c) {
    if (!values) throw std::runtime_error("null values");
    c.init_from_points(values);
  });
}

int adg_init_separable(adg_ctx* h, adg_sep_fn f, void* user, int nfunc, double eps) {
  return guard(h, [&](adg::Context& c) {
    if (!f) throw std::runtime_error("null separable function");
    c.init_separable(f, user, nfunc, eps);
  });
}

double adg_sep_eval(double x, int d, int ifunc, void* user) {
  const adg_ic_params* p = static_cast<const adg_ic_params*>(user);
  const double pi = adg::PI;
  if (p->kind == 0) {
    if (d == 0) return ifunc == 0 ? std::sin(2. * pi * (x - p->shift)) : std::cos(2. * pi * (x - p->shift));
    return ifunc == 0 ? std::cos(2. * pi * x) : std::sin(2. * pi * x);
  }
  return std::sin(2. * pi * (x - p->shift));
}

int adg_filter(adg_ctx* h, double damp_coef, const double* wave_speed, double dt, int start_level_sum) {
  return guard(h, [&](adg::Context& c) {
    if (!wave_speed) throw std::runtime_error("null wave speed");
    std::vector<double> sp(static_cast<size_t>(c.E()) * c.D);
    for (int64_t e = 0; e < c.E(); ++e)
      for (int d = 0; d < c.D; ++d) sp[e * c.D + d] = wave_speed[d];
    c.filter(damp_coef, sp.data(), dt, start_level_sum);
  });
}

int adg_filter_local(adg_ctx* h, double damp_coef, adg_coef_fn speed, void* user, double dt, int start_level_sum) {
  return guard(h, [&](adg::Context& c) {
    if (!speed) throw std::runtime_error("null speed function");
    if (c.D != 2 && c.D != 3) throw std::runtime_error("filter_local() supports dim 2 and 3");
    std::vector<double> sp(static_cast<size_t>(c.E()) * c.D, 0.);
    const int nc = 1 << c.D;
    for (int64_t e = 0; e < c.E(); ++e) {
      double xl[3], xr[3];
      for (int d = 0; d < c.D; ++d) {
        const int s = c.uslot[e * c.D + d];
        adg::cell_support(adg::level_of_slot(s), adg::suppt_of_slot(s), xl[d], xr[d]);
      }
      for (int corner = 0; corner < nc; ++corner) {  // corners in the reference's nested-loop order
        double x[3];
        for (int d = 0; d < c.D; ++d) x[d] = ((corner >> (c.D - 1 - d)) & 1) ? xr[d] : xl[d];
        for (int d = 0; d < c.D; ++d)
          sp[e * c.D + d] = std::max(std::abs(speed(x, d, user)), sp[e * c.D + d]);
      }
    }
    c.filter(damp_coef, sp.data(), dt, start_level_sum);
  });
}

int adg_eval_tensor(adg_ctx* h, int nx, const double* x1d, double* out) {
  return guard(h, [&](adg::Context& c) {
    if (!x1d || !out) throw std::runtime_error("null coordinates or output");
    c.eval_tensor(nx, x1d, out);
  });
}

int adg_error_no_separable(adg_ctx* h, adg_point_fn f, void* user, int gauss_points, double* err) {
  return guard(h, [&](adg::Context& c) {
    if (!f || !err) throw std::runtime_error("null function or output");
    if (gauss_points < 1 || gauss_points > 10) throw std::runtime_error("gauss_points must be in 1..10");
    int lmax = 0;
    for (int64_t e = 0; e < c.E(); ++e)
      for (int d = 0; d < c.D; ++d) lmax = std::max(lmax, c.level(e, d));
    const int N = 1 << lmax;
    const double h = (1. - 0.) / N;
    const std::vector<double>& qx = adg::gl_nodes(gauss_points);
    const std::vector<double>& qw = adg::gl_weights(gauss_points);
    const int g = gauss_points;
    // Quad::linear_map(x, -1, 1, tl, tr) of every 1D node of every cell
    std::vector<double> x1d(static_cast<size_t>(N) * g);
    for (int i = 0; i < N; ++i) {
      const double tl = i * h, tr = (i + 1) * h;
      for (int q = 0; q < g; ++q) x1d[i * g + q] = (qx[q] - (-1.)) / (1. - (-1.)) * (tr - tl) + tl;
    }
    const int nx = N * g;
    int64_t npts = 1;
    for (int d = 0; d < c.D; ++d) npts *= nx;
    std::vector<double> uh(static_cast<size_t>(npts));
    c.eval_tensor(nx, x1d.data(), uh.data());
    // cells, then the points of a cell, in IterativeNestedLoop order (last index fastest)
    const int D = c.D;
    int64_t ncell = 1, nq = 1;
    for (int d = 0; d < D; ++d) { ncell *= N; nq *= g; }
    std::vector<double> qwm(nq);
    for (int64_t q = 0; q < nq; ++q) {
      double w = 1;
      int64_t r = q;
      std::vector<int> qi(D);
      for (int d = D - 1; d >= 0; --d) { qi[d] = static_cast<int>(r % g); r /= g; }
      for (int d = 0; d < D; ++d) w *= qw[qi[d]];
      qwm[q] = w;
    }
    double L1 = 0, L2 = 0, Linf = 0;
    std::vector<double> x(D), av(nq);
    std::vector<int> ci(D), qi(D);
    for (int64_t cell = 0; cell < ncell; ++cell) {
      int64_t r = cell;
      for (int d = D - 1; d >= 0; --d) { ci[d] = static_cast<int>(r % N); r /= N; }
      for (int64_t q = 0; q < nq; ++q) {
        int64_t rq = q;
        for (int d = D - 1; d >= 0; --d) { qi[d] = static_cast<int>(rq % g); rq /= g; }
        int64_t idx = 0;
        for (int d = 0; d < D; ++d) {
          const int ix = ci[d] * g + qi[d];
          x[d] = x1d[ix];
          idx = idx * nx + ix;
        }
        av[q] = std::abs(std::abs(uh[idx] - f(x.data(), D, user)));
      }
      for (int64_t q = 0; q < nq; ++q) {
        L1 += qwm[q] * av[q];
        L2 += qwm[q] * av[q] * av[q];
        Linf = std::max(Linf, av[q]);
      }
    }
    L1 *= std::pow(h / 2., D);
    L2 = std::pow(L2 * std::pow(h / 2., D), 0.5);
    err[0] = L1;
    err[1] = L2;
    err[2] = Linf;
  });
}

int adg_run_linear(adg_ctx* h, const double* coef, double dt, int64_t nsteps, adg_run_stats* st) {
  return guard(h, [&](adg::Context& c) {
    c.sync();
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t n = 0; n < nsteps; ++n) {
      c.rk_init();
      for (int s = 0; s < 3; ++s) {
        c.rhs_zero();
        c.rhs_linear(coef);
        c.rk_stage(ADG_RK3, s, dt);
      }
    }
    c.sync();
    c.check_error_flag();
    const auto t1 = std::chrono::steady_clock::now();
    if (st) {
      st->steps = nsteps;
      st->dof_stages = nsteps * 3 * c.E() * c.K;
      st->max_elem = c.E();
      st->seconds = std::chrono::duration<double>(t1 - t0).count();
      st->final_time = dt * nsteps;
    }
  });
}

int adg_run_hyperbolic(adg_ctx* h, const adg_problem* p, adg_run_stats* st) {
  return guard(h, [&](adg::Context& c) {
    c.sync();
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t redo0 = c.stat_redo, proposed0 = c.stat_proposed;
    if (p->adaptive && c.refine_mode == ADG_REFINE_LEARNED && !c.prop_ready)
      throw std::runtime_error("learned refinement: the proposer has no parameters (train or set them first)");
    const int nb = c.B();
    const std::vector<double> eps =
        static_cast<int>(c.inst_eps.size()) == nb ? c.inst_eps : std::vector<double>(nb, p->eps);
    const std::vector<double> eta =
        static_cast<int>(c.inst_eta.size()) == nb ? c.inst_eta : std::vector<double>(nb, p->eta);
    std::vector<double> t(nb, p->t_start);
    std::vector<char> active(nb, 1), stashed(nb, 0);
    std::vector<std::vector<int32_t>> fin_slots(nb);
    std::vector<std::vector<double>> fin_u(nb);
    int64_t steps = 0, dof_stages = 0, max_elem = c.E();
    const adg::Context::StageProblem sp = stage_problem(p, c.D);
    auto any_active = [&]() {
      for (int i = 0; i < nb; ++i)
        if (active[i] && t[i] < p->t_final) return true;
      return false;
    };
    for (int i = 0; i < nb; ++i) active[i] = t[i] < p->t_final;
    while (any_active() && (p->max_steps <= 0 || steps < p->max_steps)) {
      // dt = cfl / sum_d |c_d| 2^{L_d} per instance (example/02_hyperbolic_05_burgers_adapt.cpp:234-244)
      adg::Context::HostTimer step_timer(c, adg::Context::HOST_STEP);
      c.inst_dt.assign(nb, 0.);
      double dt1 = 0.;
      for (int i = 0; i < nb; ++i) {
        if (!active[i]) continue;
        double sum_c_dx = 0.;
        for (int d = 0; d < c.D; ++d) {
          const int lmax = c.inst_lmax[static_cast<size_t>(i) * c.D + d];
          const double ws = c.inst_wave.empty() ? p->wave_speed[d] : c.inst_wave[static_cast<size_t>(i) * c.D + d];
          sum_c_dx += std::abs(ws) * std::pow(2., lmax);
        }
        double dt = p->cfl / sum_c_dx;
        dt = std::min(dt, p->t_final - t[i]);
        c.inst_dt[i] = dt;
        dt1 = dt;
      }
      if (nb > 1) c.upload_inst_dt();
      else c.use_inst_dt = false;
      auto stage = [&](int scheme, int s) {
        if (p->form == 1) {
          c.stage_fused(sp, scheme, s, dt1);
        } else {
          hyperbolic_rhs(c, p);
          c.rk_stage(scheme, s, dt1);
        }
      };
      if (p->adaptive && c.refine_mode != 0) {
        // proposer, RK3 step and threshold judge (redoes the instances it refines)
        c.proposer_step([&](int scheme, int s, double) { stage(scheme, s); }, eps, c.inst_dt.data());
        max_elem = std::max(max_elem, c.E());
      } else {
        if (p->adaptive) {
          std::vector<float> lf;
          std::vector<int32_t> ll;
          if (c.log_decisions) log_features(c, eps, lf, ll);
          c.copy_to_predict();
          if (p->form == 1) {
            c.rk_step(sp, ADG_RK_EULER, 1, dt1);
          } else {
            c.rk_init();
            stage(ADG_RK_EULER, 0);
          }
          if (c.log_decisions) log_labels(c, eps, lf, ll);
          c.refine_instances(eps);
          c.copy_predict_to_ucoe();
          max_elem = std::max(max_elem, c.E());
        }
        if (p->form == 1) {
          c.rk_step(sp, ADG_RK3, 3, dt1);
        } else {
          c.rk_init();
          for (int s = 0; s < 3; ++s) stage(ADG_RK3, s);
        }
      }
      dof_stages += 3 * c.E() * c.K;
      if (p->adaptive) c.coarsen_instances(eta);
      // a non-finite stage value (or a hole found by the surplus transform)
      // ends the run with an error instead of stepping on
      c.check_error_flag();
      ++steps;
      bool finished = false;
      for (int i = 0; i < nb; ++i) {
        if (!active[i]) continue;
        t[i] += c.inst_dt[i];
        if (t[i] >= p->t_final) { active[i] = 0; finished = true; }
      }
      // instances that reached t_final leave the batch while others continue
      if (finished && any_active()) {
        // keep the finished instances' elements and coefficients on the host
        std::vector<char> drop(nb, 0);
        for (int i = 0; i < nb; ++i) {
          if (active[i] || stashed[i]) continue;
          drop[i] = 1;
          stashed[i] = 1;
          fin_slots[i] = c.instance_slots(i);
          fin_u[i].resize(static_cast<size_t>(c.off[i + 1] - c.off[i]) * c.K);
          if (!fin_u[i].empty())
            adg::hip_check(hipMemcpyAsync(fin_u[i].data(), c.u.p + c.off[i] * c.K, fin_u[i].size() * sizeof(double),
                                          hipMemcpyDeviceToHost, c.stream), "download");
        }
        c.sync();
        if (c.drop_instances_device(drop) < 0) {
          c.ensure_host_grid();
          std::vector<std::vector<int32_t>> ns(nb);
          for (int i = 0; i < nb; ++i)
            if (active[i]) ns[i] = c.grids[i].slot;
          c.set_grids(ns, true);
        }
      }
    }
    // put every instance back, at its final state
    bool any_stashed = false;
    for (int i = 0; i < nb; ++i) any_stashed = any_stashed || stashed[i];
    if (any_stashed) {
      std::vector<double> uall(static_cast<size_t>(c.E()) * c.K);
      if (c.E() > 0)
        adg::hip_check(hipMemcpyAsync(uall.data(), c.u.p, uall.size() * sizeof(double), hipMemcpyDeviceToHost,
                                      c.stream), "download");
      c.sync();
      std::vector<std::vector<int32_t>> ns = c.all_instance_slots();
      std::vector<double> unew;
      for (int i = 0; i < nb; ++i) {
        if (stashed[i]) {
          ns[i] = fin_slots[i];
          unew.insert(unew.end(), fin_u[i].begin(), fin_u[i].end());
        } else {
          unew.insert(unew.end(), uall.begin() + c.off[i] * c.K, uall.begin() + c.off[i + 1] * c.K);
        }
      }
      c.set_grids_fresh(ns);
      if (!unew.empty())
        adg::hip_check(hipMemcpyAsync(c.u.p, unew.data(), unew.size() * sizeof(double), hipMemcpyHostToDevice,
                                      c.stream), "upload");
    }
    c.use_inst_dt = false;
    c.sync();
    c.check_error_flag();
    const auto t1 = std::chrono::steady_clock::now();
    if (st) {
      st->steps = steps;
      st->dof_stages = dof_stages;
      st->max_elem = max_elem;
      st->seconds = std::chrono::duration<double>(t1 - t0).count();
      double tmin = p->t_final;
      for (double v : t) tmin = std::min(tmin, v);
      st->final_time = tmin;
      st->redone = c.stat_redo - redo0;
      st->proposed = c.stat_proposed - proposed0;
    }
  }, kDeviceTables);
}

int adg_grid_init_batch(adg_ctx* h, int ninst, int sparse, int level_init) {
  return guard(h, [&](adg::Context& c) {
    if (ninst < 1) throw std::runtime_error("ninst must be >= 1");
    if (level_init < 0 || level_init > c.nmax) throw std::runtime_error("level_init must be in [0, nmax]");
    const std::vector<int32_t> s = adg::HostGrid::initial_slots(c.D, level_init, sparse != 0);
    c.set_grids(std::vector<std::vector<int32_t>>(ninst, s), false);
    c.inst_eps.clear();
    c.inst_eta.clear();
    c.sync();
  });
}

int adg_init_adaptive_intp_batch(adg_ctx* h, adg_point_fn f, void* const* users, double eps) {
  return guard(h, [&](adg::Context& c) {
    if (!f || !users) throw std::runtime_error("null point function or user array");
    c.init_adaptive_intp_batch(f, users, eps);
  });
}

int adg_num_instances(adg_ctx* h) { return (h && h->c) ? h->c->B() : -1; }

int adg_instance_range(adg_ctx* h, int inst, int64_t* first, int64_t* count) {
  return guard(h, [&](adg::Context& c) {
    if (inst < 0 || inst >= c.B()) throw std::runtime_error("instance out of range");
    if (first) *first = c.off[inst];
    if (count) *count = c.off[inst + 1] - c.off[inst];
  }, kDeviceTables);
}

int adg_set_instance_thresholds(adg_ctx* h, const double* eps, const double* eta) {
  return guard(h, [&](adg::Context& c) {
    if (!eps || !eta) throw std::runtime_error("null threshold array");
    c.inst_eps.assign(eps, eps + c.B());
    c.inst_eta.assign(eta, eta + c.B());
  });
}

int adg_set_instance_params(adg_ctx* h, const double* coef_scale, const double* alpha, const double* wave_speed) {
  return guard(h, [&](adg::Context& c) {
    const size_t n = static_cast<size_t>(c.B()) * c.D;
    for (const double* a : {coef_scale, alpha, wave_speed})
      if (a)
        for (size_t i = 0; i < n; ++i)
          if (!std::isfinite(a[i])) throw std::runtime_error("instance parameters must be finite");
    if (coef_scale) c.inst_scale.assign(coef_scale, coef_scale + n);
    else c.inst_scale.clear();
    if (alpha) c.inst_alpha.assign(alpha, alpha + n);
    else c.inst_alpha.clear();
    if (wave_speed) c.inst_wave.assign(wave_speed, wave_speed + n);
    else c.inst_wave.clear();
    if (!c.inst_alpha.empty()) c.upload_inst_alpha();
    c.vcoef_version = -1;  // the variable coefficient is sampled again
  });
}

int adg_set_kernel_variant(adg_ctx* h, int variant) {
  return guard(h, [&](adg::Context& c) {
    if (variant < -1 || variant >= adg::Context::kVariants) throw std::runtime_error("kernel variant must be -1 or 0..9");
    c.kernel_variant = variant;
  });
}

int adg_set_stage_launch(adg_ctx* h, int mode) {
  return guard(h, [&](adg::Context& c) {
    if (mode != 0 && mode != 1) throw std::runtime_error("stage launch mode must be 0 or 1");
    // The persistent cooperative stage is switched off: in some runs its
    // launch did not finish on the device, with inputs on which other runs
    // finished, and the cause is not found.
    if (mode == 1) throw std::runtime_error("persistent stage launch is switched off (did not always finish)");
    c.stage_launch = mode;
  });
}

int adg_stage_launch_count(adg_ctx* h, int64_t* launches, int64_t* stages) {
  return guard(h, [&](adg::Context& c) {
    if (launches) *launches = c.stage_launches;
    if (stages) *stages = c.stage_count;
  }, kDeviceTables);
}

int adg_set_class_variants(adg_ctx* h, const int* v, int n) {
  return guard(h, [&](adg::Context& c) {
    if (n < 0 || n > 8 || (n > 0 && !v)) throw std::runtime_error("class variants: 0..8 values expected");
    for (int i = 0; i < n; ++i)
      if (v[i] < -1 || v[i] >= adg::Context::kVariants) throw std::runtime_error("kernel variant must be -1..9");
    for (int i = 0; i < 8; ++i) c.class_variant[i] = i < n ? v[i] : -1;
  });
}

int adg_kernel_tuning(adg_ctx* h, int32_t* pairs, int cap, int* count) {
  return guard(h, [&](adg::Context& c) {
    int n = 0;
    for (const auto& kv : c.tuned) {
      if (pairs && n < cap) {
        pairs[2 * n] = kv.first;
        pairs[2 * n + 1] = kv.second;
      }
      ++n;
    }
    if (count) *count = n;
  });
}

int adg_set_graph_step(adg_ctx* h, int on) {
  return guard(h, [&](adg::Context& c) {
    if (on != 0 && on != 1) throw std::runtime_error("graph step must be 0 or 1");
    if (!on) c.drop_step_graphs();
    c.graph_step = on == 1;
  });
}

int adg_graph_step_count(adg_ctx* h, int64_t* captures, int64_t* replays) {
  return guard(h, [&](adg::Context& c) {
    if (captures) *captures = c.graph_captures;
    if (replays) *replays = c.graph_replays;
  }, kDeviceTables);
}

int adg_rk_step(adg_ctx* h, const adg_problem* p, int scheme, double dt) {
  return guard(h, [&](adg::Context& c) {
    if (!p) throw std::runtime_error("null problem");
    const int nstage = scheme == ADG_RK_EULER ? 1 : scheme == ADG_RK2 ? 2 : scheme == ADG_RK3 ? 3 : 0;
    if (nstage == 0) throw std::runtime_error("invalid RK scheme");
    c.rk_step(stage_problem(p, c.D), scheme, nstage, dt);
  }, kDeviceTables);
}

int adg_stage(adg_ctx* h, const adg_problem* p, int scheme, int stage, double dt) {
  return guard(h, [&](adg::Context& c) {
    if (!p) throw std::runtime_error("null problem");
    if (!((scheme == ADG_RK_EULER && stage == 0) || (scheme == ADG_RK2 && stage >= 0 && stage <= 1) ||
          (scheme == ADG_RK3 && stage >= 0 && stage <= 2)))
      throw std::runtime_error("invalid RK scheme/stage");
    c.stage_fused(stage_problem(p, c.D), scheme, stage, dt);
  }, kDeviceTables);
}


int adg_state_write(adg_ctx* h, const char* path) {
  return guard(h, [&](adg::Context& c) {
    if (!path) throw std::runtime_error("null path");
    std::vector<double> u(static_cast<size_t>(c.E()) * c.K);
    adg::hip_check(hipMemcpyAsync(u.data(), c.u.p, u.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                   "state download");
    c.sync();
    FilePtr f(std::fopen(path, "wb"));
    if (!f) throw std::runtime_error(std::string("state file: cannot open ") + path);
    const int32_t hdr[6] = {c.D, c.k, c.m, c.nmax, c.B(), 0};
    write_all(f.get(), kStateMagic, sizeof(kStateMagic));
    write_all(f.get(), hdr, sizeof(hdr));
    for (int i = 0; i < c.B(); ++i) {
      const int64_t n = c.off[i + 1] - c.off[i];
      write_all(f.get(), &n, sizeof(n));
    }
    std::vector<int32_t> ls(static_cast<size_t>(c.E()) * 2 * c.D);
    for (int64_t e = 0; e < c.E(); ++e)
      for (int d = 0; d < c.D; ++d) {
        ls[(e * 2) * c.D + d] = c.level(e, d);
        ls[(e * 2 + 1) * c.D + d] = adg::suppt_of_slot(c.uslot[e * c.D + d]);
      }
    write_all(f.get(), ls.data(), ls.size() * sizeof(int32_t));
    write_all(f.get(), u.data(), u.size() * sizeof(double));
    if (std::fflush(f.get()) != 0) throw std::runtime_error("state file: write failed");
  });
}

int adg_state_read(adg_ctx* h, const char* path) {
  return guard(h, [&](adg::Context& c) {
    if (!path) throw std::runtime_error("null path");
    FilePtr f(std::fopen(path, "rb"));
    if (!f) throw std::runtime_error(std::string("state file: cannot open ") + path);
    char magic[8];
    int32_t hdr[6];
    read_all(f.get(), magic, sizeof(magic));
    if (std::memcmp(magic, kStateMagic, sizeof(magic)) != 0) throw std::runtime_error("state file: bad magic");
    read_all(f.get(), hdr, sizeof(hdr));
    if (hdr[0] != c.D || hdr[1] != c.k || hdr[2] != c.m || hdr[3] != c.nmax)
      throw std::runtime_error("state file: dimension, degrees or maximum level differ from the context");
    const int B = hdr[4];
    if (B < 1 || B > (1 << 20)) throw std::runtime_error("state file: bad instance count");
    std::vector<int64_t> cnt(B);
    read_all(f.get(), cnt.data(), cnt.size() * sizeof(int64_t));
    std::vector<std::vector<int32_t>> slots(B);
    int64_t tot = 0;
    for (int i = 0; i < B; ++i) {
      if (cnt[i] < 1 || cnt[i] > (int64_t(1) << 31)) throw std::runtime_error("state file: bad element count");
      std::vector<int32_t> ls(static_cast<size_t>(cnt[i]) * 2 * c.D);
      read_all(f.get(), ls.data(), ls.size() * sizeof(int32_t));
      slots[i].resize(static_cast<size_t>(cnt[i]) * c.D);
      for (int64_t e = 0; e < cnt[i]; ++e)
        for (int d = 0; d < c.D; ++d) {
          const int32_t l = ls[(e * 2) * c.D + d], j = ls[(e * 2 + 1) * c.D + d];
          if (l < 0 || l > c.nmax) throw std::runtime_error("state file: level out of range");
          const int jmax = l == 0 ? 1 : (1 << l) - 1;
          if (j < 1 || j > jmax || j % 2 == 0) throw std::runtime_error("state file: support index out of range");
          slots[i][e * c.D + d] = adg::slot_of(l, j);
        }
      tot += cnt[i];
    }
    std::vector<double> u(static_cast<size_t>(tot) * c.K);
    read_all(f.get(), u.data(), u.size() * sizeof(double));
    char extra;
    if (std::fread(&extra, 1, 1, f.get()) != 0) throw std::runtime_error("state file: trailing bytes");
    // the file's element order is hash-key order, which the grid rebuild restores;
    // map the coefficients by key
    std::vector<std::vector<int32_t>> keyed = slots;
    c.set_grids(slots, false);
    c.inst_eps.clear();
    c.inst_eta.clear();
    if (c.E() != tot) throw std::runtime_error("state file: duplicate elements");
    std::vector<double> ud(u.size());
    int64_t src = 0;
    for (int i = 0; i < B; ++i) {
      std::unordered_map<int32_t, int64_t> pos;
      for (int64_t e = c.off[i]; e < c.off[i + 1]; ++e) pos[c.uhash[e]] = e;
      for (int64_t e = 0; e < cnt[i]; ++e, ++src) {
        int32_t l[adg::kMaxDim], j[adg::kMaxDim];
        for (int d = 0; d < c.D; ++d) {
          l[d] = adg::level_of_slot(keyed[i][e * c.D + d]);
          j[d] = adg::suppt_of_slot(keyed[i][e * c.D + d]);
        }
        auto it = pos.find(c.hk.key(l, j, c.D));
        if (it == pos.end()) throw std::runtime_error("state file: element lost in the grid rebuild");
        std::copy(&u[src * c.K], &u[src * c.K] + c.K, &ud[it->second * c.K]);
      }
    }
    adg::hip_check(hipMemcpyAsync(c.u.p, ud.data(), ud.size() * sizeof(double), hipMemcpyHostToDevice, c.stream),
                   "state upload");
    c.sync();
  });
}

int adg_state_save(adg_ctx* h, int slot) { return guard(h, [&](adg::Context& c) { c.save_state(slot); }); }
int adg_state_restore(adg_ctx* h, int slot) {
  return guard(h, [&](adg::Context& c) { c.restore_state(slot); }, kDeviceTables);
}

int adg_profile(adg_ctx* h, int enable) {
  return guard(h, [&](adg::Context& c) {
    c.prof_reset();
    c.prof_on = enable != 0;
  }, kDeviceTables);
}

int adg_profile_host(adg_ctx* h, double* ms) {
  return guard(h, [&](adg::Context& c) {
    for (int i = 0; i < adg::Context::HOST_NCAT; ++i)
      if (ms) ms[i] = c.host_ms[i];
  }, kDeviceTables);
}

int adg_profile_get(adg_ctx* h, double* ms, int64_t* count, double* bytes) {
  return guard(h, [&](adg::Context& c) {
    c.prof_collect();
    for (int i = 0; i < adg::Context::PROF_NCAT; ++i) {
      if (ms) ms[i] = c.prof_ms[i];
      if (count) count[i] = c.prof_count[i];
      if (bytes) bytes[i] = c.prof_bytes[i];
    }
  }, kDeviceTables);
}

}  // extern "C"
