// Device context of one adaptive sparse-grid DG solution (internal C++ API
// behind the C-ABI in include/adg.h).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <functional>
#include <map>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "grid.hpp"
#include "setup.hpp"

namespace adg {

struct TaskDev;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);
// runs fn(0..n-1) on the library's host thread pool (serially for small n)
void parallel_for(int n, const std::function<void(int)>& fn);

// Allocation-free entry points (noalloc.hip): capacity of the per-GPU
// scratch array and its device address on a device (set up by
// noalloc_setup with that device current, once per device).
constexpr size_t kNoallocScratchDoubles = size_t{1} << 22;
void noalloc_setup(int device);
double* noalloc_scratch(int device);

constexpr int ADG_FLUX_VARCOEF_ID = 2;         // ADG_FLUX_VARCOEF of include/adg.h
constexpr int ADG_FLUX_VARCOEF_SCALED_ID = 3;  // ADG_FLUX_VARCOEF_SCALED
inline bool is_varcoef(int kind) { return kind == ADG_FLUX_VARCOEF_ID || kind == ADG_FLUX_VARCOEF_SCALED_ID; }

// Growable device buffer.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  void swap(DevBuf& o) noexcept {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
  void ensure(size_t n) {
    if (n <= cap) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    size_t c = cap ? cap : 1024;
    while (c < n) c *= 2;
    hip_check(hipMalloc(&p, c * sizeof(T)), "hipMalloc");
    cap = c;
  }
};

// Learned refinement proposer: a multilayer perceptron with kPropIn inputs,
// two hidden layers of kPropHidden ReLU units and one logistic output, the
// probability that the plain predict-and-refine step would give a leaf
// element its children.  Parameters (FP32, row-major): W1[H][F], b1[H],
// W2[H][H], b2[H], W3[H], b3.
constexpr int kPropIn = 12;
constexpr int kPropHidden = 64;
constexpr int kPropParams =
    kPropHidden * kPropIn + kPropHidden + kPropHidden * kPropHidden + kPropHidden + kPropHidden + 1;

// A 1D operator table on the device, stored target-major in cell blocks:
// blk[((st * ns + s) * pf1 + k) * pt1 + i] = M(row s*pf1 + k, column st*pt1 + i)
// of the dense level-ordered table (row = source basis, column = target basis),
// so the sources of one target cell are read from consecutive blocks.
struct SweepTable {
  const double* dev = nullptr;
  int ns = 0;  // 1D cells (2^NMAX)
};

class Context {
 public:
  Context(int dim, int k, int m, int nmax);
  ~Context();

  const int D, k, m, nmax;
  int K, P;  // (k+1)^D, (m+1)^D

  // host state
  Basis1D alpt, lagr;
  HashKey hk;
  // One element table per independent solve ("instance"); the device holds
  // their union, instance after instance.  Fibers never cross instances, so
  // every kernel serves all instances in one launch.
  std::vector<HostGrid> grids;
  std::vector<int64_t> off{0};          // union offset of each instance, size B+1
  std::vector<int32_t> uslot, uinst;    // union: slots [E*D], instance [E]
  std::vector<int32_t> uhash;           // union: hash keys [E]
  int B() const { return static_cast<int>(grids.size()); }
  int level(int64_t e, int d) const { return level_of_slot(uslot[e * D + d]); }
  // per-instance adaptivity thresholds set by the caller (empty: the run's
  // scalar eps / eta apply to every instance) and per-instance step sizes
  std::vector<double> inst_eps, inst_eta, inst_dt;
  // per-instance problem parameters (B x D each; empty: the problem's own):
  // scale of the variable coefficient (ADG_FLUX_VARCOEF_SCALED), LF
  // constants and wave speeds of the time-step rule
  std::vector<double> inst_scale, inst_alpha, inst_wave;
  DevBuf<double> d_inst_dt, d_pen_coef, d_inst_scale;  // d_pen_coef[d * B + i] = -alpha_{i,d} / 2
  DevBuf<double> vcoef_1d;  // 1 + s sin(2 pi x) at the points of a 1D grid
  int64_t vcoef_version = -1;
  // uploads d_pen_coef from inst_alpha (per-instance LF penalty constants)
  void upload_inst_alpha();
  DevBuf<int32_t> g_inst;
  bool use_inst_dt = false;
  std::vector<double> up_host;    // point values of the adaptive interpolation
  std::vector<char> new_add;      // Element::new_add of the adaptive interpolation
  std::vector<double> lin_coef;   // coefficients of the uploaded linear tables
  int flux_kind = -1;
  double flux_param = 0.;

  hipStream_t stream = nullptr;
  int device = 0;  // the GPU current at creation (the context's stream lives there)

  // tables
  DevBuf<double> t_pt, t_la_uv, t_la_uvx, t_la_flx, t_aa[5], t_lin[kMaxDim], t_lmass, t_wt;
  DevBuf<int32_t> t_anc_slot, t_anc_point;
  int na = 0, nl = 0;  // 1D basis counts

  // Interpolation family of eval_up / flux / fp_to_coe / rhs_lagr and the
  // fused stage: 0 Lagrange, 1 Hermite P3 (HermInterpolation with
  // HyperbolicHermRHS).  The family in use has its point, operator and
  // surplus tables in t_pt, t_la_*, t_wt, t_anc_*; the other family's wait in
  // alt_*.  The Hermite tables are built on first use.
  int interp = 0;
  std::unique_ptr<Basis1D> herm;
  DevBuf<double> alt_pt, alt_uv, alt_uvx, alt_flx, alt_wt;
  DevBuf<int32_t> alt_anc_slot, alt_anc_point;
  void set_interpolation(int family);
  const Basis1D& interp_basis() const { return interp ? *herm : lagr; }
  // The adaptive interpolation initialisation always runs on the Lagrange
  // family (as DGAdaptIntp does); this restores the caller's family after it.
  struct LagrangeScope {
    Context& c;
    int saved;
    explicit LagrangeScope(Context& ctx) : c(ctx), saved(ctx.interp) { c.set_interpolation(0); }
    ~LagrangeScope() { c.set_interpolation(saved); }
  };

  // grid arrays
  DevBuf<int32_t> g_slot[kMaxDim], g_fbase[kMaxDim], g_fmaxl[kMaxDim], g_fmap[kMaxDim];
  DevBuf<uint32_t> g_fbits[kMaxDim];  // occupancy bit of every fiber-map entry (+ one zero word)
  int64_t fbits_words[kMaxDim] = {0, 0, 0, 0};
  int maxfib = 1;  // largest fiber (elements) over all dimensions and instances
  // fiber order of every dimension for the fiber-tile kernels (built lazily
  // after a grid change): elements by (fiber, slot) and prefix popcounts of
  // the occupancy words
  DevBuf<int32_t> g_perm[kMaxDim], g_pre[kMaxDim], g_cnt;
  DevBuf<uint8_t> scan_tmp;
  int64_t order_version = -1;
  void ensure_fiber_order();
  // tiles of whole fibers, te = 64 << which elements per tile window
  struct TileSetH {
    int te = 0, max_tile = 0;
    int64_t version = -1, ntile = 0;
    DevBuf<int32_t> tbeg[kMaxDim], tend[kMaxDim];
  };
  TileSetH tsets[3];
  TileSetH& tile_set(int which);
  template <class T>
  void tile_launch(int variant, int pf1, int pt1, const struct TaskDev* tasks, int R, int ntask, bool last);
  DevBuf<double> tile_tv, tile_pv;  // per-term values of the tile form's last sweeps
  DevBuf<int32_t> g_remap;
  DevBuf<int32_t> g_lsum;  // level sum of each element
  // Level-sum cap of the reference's _coarse_grid variants: sweeps and
  // operators ignore elements above it (their output is zero).  -1: none.
  int lsum_cap = -1;

  // fields
  DevBuf<double> u, tn, pred, rhs, up, uintp, fp, fucoe, interA, interB, tmpA, tmpB, vcoef, sumsq, remap_tmp;
  DevBuf<int32_t> err_flag;
  bool same_flux = false;  // all flux components equal (Burgers): one surplus transform

  int64_t E() const { return off.back(); }

  void set_grid_slots(const std::vector<int32_t>& slots, bool keep_fields);
  // Replace all instance grids; keep_fields remaps u and the predict copy.
  // changed (optional): instances flagged 0 keep their current table.
  void set_grids(const std::vector<std::vector<int32_t>>& slots, bool keep_fields,
                 const std::vector<char>* changed = nullptr);
  void upload_grid();
  void ensure_fields();

  // operators (stream ordered)
  void sweep(const double* in, const int* size_in, double* out, double* accum, const int* size_out, int dd,
             SweepTable tab, bool flx, int lu, double coef, int pf, int pt, const double* coef_inst = nullptr);
  void partial_sums(const SweepTable* tabs, const int* flx, int pf, int pt, double coef, const double* src, double* dst);
  void eval_up();
  void flux(int kind, const double* params);
  void fp_to_coe();
  void rhs_zero();
  void rhs_lagr(bool flux_term);
  void rhs_single_dim(SweepTable tab, int d, double coef, const double* coef_inst = nullptr);
  void rhs_penalty(const double* alpha);
  void rhs_linear(const double* coef);
  void rk_init();
  void rk_stage(int scheme, int stage, double dt);
  void copy_to_predict();
  void copy_predict_to_ucoe();
  void indicator_sumsq(double* host_out);
  void indicator_sumsq_dev();  // sums of squares of every element into sumsq
  int64_t refine(double eps);
  int64_t coarsen(double eta);
  // one threshold per instance
  int64_t refine_instances(const std::vector<double>& eps);
  int64_t coarsen_instances(const std::vector<double>& eta);
  // Refine / coarsen on the device (adapt_dev.hip): threshold tests, child
  // generation, hole closure, hash keys, the sorted union table and its fiber
  // tables are computed on the GPU and the host reads back counts only.  The
  // host copy of the tables (grids, uslot, uhash, uinst) is then stale until
  // ensure_host_grid() rebuilds it from the device slots.  A call returns -1
  // with nothing changed when the host has to decide instead: a norm within
  // a few ulps of its threshold (where the reference's pow decides), a full
  // table, or sizes whose keys do not pack into 63 bits.
  bool device_adapt = true;  // (adg_set_device_adapt; the host form with 0)
  bool host_stale = false;
  int64_t stat_dev_adapt = 0, stat_dev_fallback = 0;
  int32_t dev_fallback_flags = 0;  // why calls went to the host: 1 norm at a threshold, 2 table full
  int64_t refine_device(const std::vector<double>& eps);
  // refine on the device the leaves flagged in flags (device array, one per element)
  int64_t refine_device_flags(const int32_t* flags);
  int64_t refine_device_impl(const std::vector<double>* eps, const int32_t* flags);
  // flags[e] = (mask[e] if mask) && norm of e above thr of its instance, on
  // the device; norms within the band where pow decides are decided on the host
  void norm_flags_dev(const std::vector<double>& thr, const int32_t* mask, int32_t* flags);
  // proposer_step with a deferring judge, all on the device
  int64_t proposer_step_device(const std::function<void(int, int, double)>& stage, const std::vector<double>& eps);
  DevBuf<int32_t> ad_leaf[2], ad_flag, ad_band;
  int64_t coarsen_device(const std::vector<double>& eta);
  void ensure_host_grid();
  // removes every element of the flagged instances on the device (they stay
  // in the batch with no elements); -1 when the host has to do it
  int64_t drop_instances_device(const std::vector<char>& drop);
  // second half of a device refine / coarsen / drop: fiber tables of the new
  // union table in ad_*, read-back of its counts (the call waits for the
  // stream), fiber maps, swap into g_*, field remap
  void device_grid_finish(int64_t N, int64_t old_E);
  bool device_adapt_fits(int64_t n) const;
  // element slots of instance i in table order (from the device when the
  // host copy is stale)
  std::vector<int32_t> instance_slots(int i);
  std::vector<std::vector<int32_t>> all_instance_slots();
  // set_grids(slots, false) with the table built on the device when device
  // adapt is on (fields zero; the host copy is left stale); false when it
  // fell back to set_grids
  bool set_grids_fresh(const std::vector<std::vector<int32_t>>& slots);
  // entries of the device tables (slots, instances, level sums, fiber bases,
  // fiber levels, fiber maps) that differ from a host build of the same grid
  int64_t check_device_grid();
  std::vector<int32_t> inst_lmax;  // [B * D] deepest level per instance and dimension
  void host_union();               // uslot, uhash, uinst and inst_lmax from grids
  DevBuf<uint64_t> g_ukey;  // (instance << 32 | hash key) of every element, ascending
  DevBuf<uint64_t> ad_hkey, ad_newkey, ad_keys[2], ad_ukey;
  DevBuf<int32_t> ad_ctr, ad_vals[2], ad_run, ad_runmax, ad_runlen, ad_meta, ad_nchild, ad_zmark, ad_dec, ad_cand,
      ad_lvtab;
  DevBuf<int32_t> ad_slot[kMaxDim], ad_fbase[kMaxDim], ad_fmaxl[kMaxDim], ad_inst, ad_lsum;
  DevBuf<double> ad_thr;
  DevBuf<uint8_t> ad_tmp;
  void init_adaptive_intp(double (*f)(const double*, int, void*), void* user, double eps);
  void init_adaptive_intp_batch(double (*f)(const double*, int, void*), void* const* users, double eps);
  void init_separable(double (*f)(double, int, int, void*), void* user, int nfunc, double eps);
  void init_from_points(const double* values);
  // u at every point of the tensor grid x1d^D (first dimension slowest)
  void eval_tensor(int nx, const double* x1d, double* out_host);
  DevBuf<uint64_t> eval_hkey;
  DevBuf<int32_t> eval_hval, eval_ncand, eval_cslot;
  DevBuf<double> eval_cval, eval_out;
  uint64_t eval_hmask = 0;
  int64_t eval_version = -1;
  void filter(double damp, const double* speed, double dt, int start);
  void upload_inst_dt();
  void check_error_flag();

  SweepTable table(int id);
  void sync();

  // Fused RK stage: 3*dim launches per stage (first sweeps of every partial
  // sum batched per phase; last sweeps, point flux, penalty and RK update
  // fused into per-output kernels).  Bit-identical to the per-sweep path.
  struct StageProblem {
    int flux_kind = 0;
    double flux_param = 0.;
    double alpha[kMaxDim] = {0, 0, 0, 0};
  };
  void stage_fused(const StageProblem& p, int scheme, int stage, double dt);
  struct FusedPlan;
  FusedPlan& fused_plan(const StageProblem& p);
  template <class T>
  void stage_body(const StageProblem& p, FusedPlan& pl, int scheme, int stage, double dt);
  // How a fused stage is launched: 0 one launch per phase (thread mappings
  // chosen by the autotuner), 1 one persistent cooperative launch per stage
  // (fiber tiles of 64 elements, four rest indices per thread; Lagrange family)
  int stage_launch = 0;
  template <class T>
  void stage_coop(const StageProblem& p, FusedPlan& pl, int scheme, int stage, double dt);
  int64_t stage_launches = 0, stage_count = 0;  // kernel launches and fused stages (counters)

  // One explicit RK step of fused stages (u^n <- u, then `nstage` stages).
  // With graph_step on, the step is recorded into a HIP graph the second
  // time it runs on the same grid (the first run builds the plans and tunes
  // the kernels) and replayed after that.  A graph holds the device pointers
  // and launch shapes of its grid, so any change of the grid, of a buffer,
  // of the problem or of the launch settings records a new one; four graphs
  // are kept (a step of each kind from each buffer of the u / u_next pair).
  void rk_step(const StageProblem& p, int scheme, int nstage, double dt);
  void rk_step_eager(const StageProblem& p, int scheme, int nstage, double dt);
  struct StepGraph {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
    int64_t launches = 0;  // kernel launches of the recorded step
  };
  std::vector<uintptr_t> step_key(const StageProblem& p, int scheme, int nstage, double dt) const;
  void drop_step_graphs();
  // Lazily built state (plans, fiber orders, tile sets, split tables, point
  // coefficients, kernel tuning) uploads host data or times launches, which a
  // graph cannot hold: while a step is recorded, such a build throws
  // NeedsBuild before it changes anything, and the step runs without a graph.
  struct NeedsBuild : std::logic_error {
    explicit NeedsBuild(const char* what) : std::logic_error(std::string("needs a build while recording: ") + what) {}
  };
  bool recording = false;
  void not_recording(const char* what) const;
  bool graph_step = false;
  static constexpr size_t kStepGraphs = 4;
  StepGraph step_graphs[kStepGraphs];
  size_t next_graph = 0;
  std::vector<std::vector<uintptr_t>> step_seen;  // states of the last steps run without a graph
  int64_t graph_captures = 0, graph_replays = 0;

  // Storage precision of the fused stage: 0 FP64; 1 FP32 fiber transforms,
  // point values, flux and surplus transform, with every 1D table and surplus
  // weight split into two FP32 parts (M = hi + lo).  The RK combination, the
  // accumulation of the right-hand side, the LF penalty and the refine /
  // coarsen norms stay FP64.  The per-sweep path is always FP64.
  int precision = 0;
  DevBuf<float> up32, fp32, fucoe32, interA32, interB32;
  struct SplitTable { DevBuf<float> hi, lo; };
  std::map<const double*, std::unique_ptr<SplitTable>> split_tabs;
  const SplitTable& split_table(const double* dev, size_t n);
  // thread mapping of the fused kernels: 0 per output, 1 per (element, rest
  // index), 2 per (element, 4 rest indices), 3 per (element, 2 rest indices),
  // 4..6 fiber tiles of 64, 128, 256 elements (LDS-staged), a thread per
  // (target, rest index); 7..9 the same tiles, a thread per target; -1 lets
  // the autotuner choose per kernel class and grid-size class
  static constexpr int kVariants = 10;
  enum { TUNE_UP_PHASE = 0, TUNE_UP_FINAL, TUNE_RHS_PHASE, TUNE_RHS_FINAL };
  int kernel_variant = -1;
  // per kernel class (0..3 FP64, 4..7 FP32): a fixed mapping, -1 the autotuner
  int class_variant[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  std::map<int, int> tuned;
  void tuned_launch(int cls, const std::function<void(int)>& launch);
  std::unique_ptr<FusedPlan> plans[2];
  DevBuf<double> u_next, fused_tmp;
  int64_t grid_version = 0;

  // ---- artificial viscosity, matrix-free (viscosity_option 3 of the
  // reference's Burgers driver) ----
  // DGAdapt::update_viscosity_element: childless elements at the finest level
  // whose norm exceeds 2^((1 - k) |l|_1) by more than kappa decades
  int64_t viscosity_elements(double kappa);
  std::vector<int64_t> visc_elems;  // union element indices at the time of the update
  // the same elements by instance and packed slot vector: they stay valid when
  // refinement renumbers the elements (the reference holds Element pointers)
  std::vector<std::vector<uint64_t>> visc_keys;
  // LargViscosInterpolation::support_gradu_Lagr_fast on the Lagrange family:
  // fucoe[d] = surplus transform of coef * du/dx_d at the points inside a
  // viscosity element's support (counted with multiplicity when overlap)
  void viscosity_flux(double coef, bool overlap);
  void eval_der_up(int d0);  // FastLagrIntp::eval_der_up_Lagr
  DevBuf<double> t_pt_d1, visc_w;

  // ---- artificial viscosity as an assembled operator (options 1 and 2) and
  // IMEX43 (imex.hip) ----
  // ArtificialViscosity::assemble_matrix over the current viscosity elements;
  // returns the number of K x K blocks
  int64_t visc_operator_assemble(double coef, int option);
  void visc_operator_check() const;
  // y = A x (mode 0), y += (A x) (mode 1), y = (I - c A) x (mode 2); device pointers
  void visc_operator_apply(const double* x, double* y, int mode, double c);
  void rhs_visc_operator();  // rhs += (A u)
  void imex43_init(double dt);
  int imex43_stage(int stage);  // largest CG iteration count, -1 without convergence
  int cg_solve(double c);
  int64_t vop_blocks = 0, vop_version = -1, vop_nnz_entries = 0, imex_version = -1;
  std::vector<int32_t> vop_key_tid;   // 1D integral key -> table entry
  std::vector<double> vop_tab_host;   // per entry: mass, stiffness ((k+1)^2 each)
  DevBuf<int32_t> vop_rowptr, vop_col, vop_recptr, vop_rec, cg_state, cg_inst_chunk;
  DevBuf<double> vop_tab, vop_val, imex_tn, imex_G[4], imex_F[2], cg_b, cg_r, cg_p, cg_t, cg_inv, cg_part, cg_scal;
  DevBuf<int64_t> cg_chunk;
  int cg_nchunk = 0;
  std::vector<int> cg_iters;  // per instance, last solve
  double imex_dt = 0.;

  // ---- learned refinement proposer (proposer.hip) ----
  // Features of every element for the proposer, from the current u:
  // [0] log10(norm / eps) / 4 of the element, [1 + 2d], [2 + 2d] the same for
  // its same-level neighbours at j - 2 and j + 2 in dimension d (periodic; a
  // missing neighbour has norm 0, levels 0 and 1 are their own neighbours),
  // [1 + 2D] sum of levels / (D nmax), [2 + 2D] max level / nmax, then zeros;
  // clamped to [-4, 4].  leaf: fewer children present than possible.
  void proposer_features(const std::vector<double>& eps, float* feats_dev, int32_t* leaf_dev);
  DevBuf<float> prop_params, prop_grad, prop_m1, prop_m2, prop_norm, prop_feats, prop_prob;
  DevBuf<float> prop_h1, prop_h2, prop_dh1, prop_dh2, prop_dz, prop_x, prop_y, prop_w, prop_loss;
  DevBuf<uint16_t> prop_params16;  // FP16 copy for inference
  DevBuf<int32_t> prop_leaf, prop_idx;
  DevBuf<double> prop_eps;
  bool prop_ready = false, prop16_valid = false;
  int64_t prop_adam_t = 0;
  void proposer_init(uint64_t seed);
  void proposer_set_params(const float* host);
  void proposer_get_params(float* host);
  // gradient of the weighted logistic loss over n samples (host arrays)
  double proposer_gradient(const float* feats, const uint8_t* labels, const float* weights, int64_t n, float* grad);
  // Adam on minibatches; returns the mean weighted loss of the last epoch
  double proposer_train(const float* feats, const uint8_t* labels, int64_t n, int epochs, int batch, float lr,
                        uint64_t seed);
  void proposer_predict(const float* feats_dev, int64_t n, float* prob_dev);  // FP16 inference
  // refine modes of the adaptive step: 0 predict by forward Euler and refine
  // (the reference), 1 learned proposer + threshold judge, 2 current norms +
  // threshold judge
  int refine_mode = 0;
  // judge of the proposer modes: 0 redo the instances it refines from u^n,
  // 1 give the missed leaves their children after the step and go on
  int judge_defer = 0;
  int64_t stat_deferred = 0;  // leaves refined by a deferring judge
  double prop_threshold = 0.5;
  int64_t stat_redo = 0, stat_proposed = 0;
  // decision log of the predict-and-refine step: features of every leaf at
  // the start of a step and whether the step refined it
  bool log_decisions = false;
  std::vector<float> log_feats;
  std::vector<uint8_t> log_labels;
  // one adaptive step with a proposer (modes 1, 2): propose, refine, RK3,
  // judge (leaves above eps without children are refined and their
  // instances redo the step from u^n); returns the number of redone steps
  int64_t proposer_step(const std::function<void(int, int, double)>& stage, const std::vector<double>& eps,
                        double* dt_inst);
  DevBuf<int32_t> inst_mask;

  // In-memory snapshots of the solution state (grid + Alpert coefficients).
  struct Snapshot { std::vector<std::vector<int32_t>> slots; DevBuf<double> u; int64_t E = 0; };
  Snapshot snaps[4];
  void save_state(int slot);
  void restore_state(int slot);

  // Kernel timing with HIP events on this context's stream.  Categories:
  // 0 sweep, 1 surplus transform, 2 flux, 3 RK stage, 4 other.
  enum { PROF_SWEEP = 0, PROF_SURPLUS, PROF_FLUX, PROF_RK, PROF_OTHER, PROF_RHS_PHASE, PROF_NCAT };
  bool prof_on = false;
  struct ProfRec { hipEvent_t a, b; int cat; double bytes; };
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  // host wall time (inclusive) of the adaptive bookkeeping, recorded while profiling
  enum { HOST_ADAPT = 0, HOST_GRID, HOST_PLAN, HOST_STEP, HOST_BUILD, HOST_UPLOAD, HOST_NCAT };
  double host_ms[HOST_NCAT] = {0};
  struct HostTimer {
    Context& c;
    int cat;
    std::chrono::steady_clock::time_point t0;
    HostTimer(Context& ctx, int category) : c(ctx), cat(category), t0(std::chrono::steady_clock::now()) {}
    ~HostTimer() {
      if (c.prof_on)
        c.host_ms[cat] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
  };
  double prof_ms[PROF_NCAT] = {0}, prof_bytes[PROF_NCAT] = {0};
  int64_t prof_count[PROF_NCAT] = {0};
  void prof_begin(int cat, double bytes);
  void prof_end();
  void prof_collect();
  void prof_reset();

 private:
  void remap_field(DevBuf<double>& f, int per_elem, const std::vector<int32_t>& old_of_new, int64_t old_E);
  // the same with old_of_new already in g_remap (new_E entries)
  void remap_field_dev(DevBuf<double>& f, int per_elem, int64_t new_E, int64_t old_E);

  void surplus_transform(const double* in, double* out, double* scratch_a, double* scratch_b, int ncomp);
  void surplus_transform32(const float* in, float* out, float* scratch_a, float* scratch_b, int ncomp);
  void compute_vcoef();
};

std::vector<std::vector<int>> transform_orders(int D, std::vector<std::vector<int>>& lus);

}  // namespace adg
