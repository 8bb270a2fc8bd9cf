/* C-ABI of the MI355X adaptive sparse-grid DG library (libadg_hip.so).
 *
 * One context holds one DG solution (DGSolution / DGAdapt state) on one GPU:
 * the sorted active-element table, the fiber maps, the 1D operator tables and
 * the coefficient arrays, all resident in device memory.  Host arrays that
 * cross this interface are borrowed for the duration of the call and are
 * ordered by element hash key (ascending); per-element tensors are row-major
 * (last dimension fastest), as VecMultiD stores them (include/VecMultiD.h:114-152
 * in the reference).
 *
 * Every function returns 0 on success and a negative ADG_E_* code on error;
 * adg_last_error() returns the message.  Calls on one context are ordered on
 * that context's HIP stream; functions that hand results to the host
 * synchronise that stream.  Contexts are independent; a context must not be
 * used from two host threads at once.
 *
 * Each entry point names the reference interface it replaces
 * (file:line in JuntaoHuang/adaptive-multiresolution-DG).
 */
#ifndef ADG_H
#define ADG_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct adg_ctx adg_ctx;
/* Initial-condition / point function: value at x[0..dim-1]. */
typedef double (*adg_point_fn)(const double* x, int dim, void* user);

enum {
  ADG_OK = 0,
  ADG_E_ARG = -1,      /* invalid argument or state */
  ADG_E_HIP = -2,      /* HIP runtime error */
  ADG_E_ANCESTOR = -3, /* surplus transform found a missing ancestor (grid has a hole) */
  ADG_E_NONFINITE = -4,/* non-finite value detected */
  ADG_E_INTERNAL = -5,
  ADG_E_NOCONVERGE = -6 /* an implicit solve ended above its tolerance */
};

/* Fields (coefficient arrays) of a context. */
enum {
  ADG_FIELD_UCOE = 0,   /* Element::ucoe_alpt, (k+1)^dim per element */
  ADG_FIELD_RHS = 1,    /* Element::rhs */
  ADG_FIELD_UP = 2,     /* Element::up_intp, values at Lagrange points, (m+1)^dim */
  ADG_FIELD_TN = 3,     /* RK stage base u^n (ExplicitRK::ucoe_tn) */
  ADG_FIELD_UINTP = 4,  /* Element::ucoe_intp (Lagrange surplus of the initial function) */
  ADG_FIELD_FUCOE = 10, /* +d: Element::fucoe_intp[0][d], Lagrange surplus of flux d */
  ADG_FIELD_FP = 20     /* +d: Element::fp_intp[0][d], flux d at the interpolation points (FP64 form) */
};

/* 1D operator tables (row = "from" basis, column = "to" basis). */
enum {
  ADG_TAB_ALPT_AT_LAGR = 0, /* Alpert basis at Lagrange points (FastLagrIntp::alpt_basis_Lagr_pt) */
  ADG_TAB_LA_UV = 1,        /* OperatorMatrix1D<Lagr,Alpt>::u_v */
  ADG_TAB_LA_UVX = 2,       /* OperatorMatrix1D<Lagr,Alpt>::u_vx */
  ADG_TAB_LA_FLX = 3,       /* ulft_vjp + urgt_vjp (Lagr,Alpt) */
  ADG_TAB_AA_UV = 4, ADG_TAB_AA_UVX = 5, ADG_TAB_AA_ULFT = 6, ADG_TAB_AA_URGT = 7,
  ADG_TAB_AA_UJP = 8        /* OperatorMatrix1D<Alpt,Alpt>::ujp_vjp */
};

/* L/U split of a 1D sweep (FastMultiplyLU::transform_1D LU argument). */
enum { ADG_LU_FULL = 0, ADG_LU_L = 1, ADG_LU_U = 2 };

/* Point flux kinds (LagrInterpolation::eval_fp_Lag / eval_coe_u_Lag). */
enum {
  ADG_FLUX_BURGERS = 0,  /* f_d(u) = u*u*0.5 for every d (FluxFunction::burgers_flux_scalar) */
  ADG_FLUX_SINCOS = 1,   /* f_0 = sin(u)*s, f_1 = cos(u)*s (FluxFunction::sin_flux/cos_flux) */
  ADG_FLUX_VARCOEF = 2,  /* f_d = a_d(x)*u, a_d(x) = 1 + s*sin(2*pi*x_{(d+1) mod dim}) */
  /* f_d = c_d a_d(x) u with the instance's scale c_d (adg_set_instance_params; 1 when unset):
   * LagrInterpolation::var_coeff_u_Lagr_fast (source/Interplation.cpp:4202-4213) with
   * coe_func(x, d) = c_d (1 + s sin(2 pi x_{(d+1) mod dim})) */
  ADG_FLUX_VARCOEF_SCALED = 3
};

/* RK schemes (ODESolver.h): forward Euler, RK2 midpoint and SSP-RK3. */
enum { ADG_RK_EULER = 0, ADG_RK2 = 2, ADG_RK3 = 3 };

/* Context: statics of the reference (AlptBasis::PMAX = k, LagrBasis::PMAX = m,
 * LagrBasis::msh_case = 1, Element::DIM = dim, DGSolution NMAX = nmax) plus the
 * AllBasis / OperatorMatrix1D / LagrInterpolation tables built on the host and
 * uploaded once.  Replaces the setup block of the example drivers
 * (example/02_hyperbolic_05_burgers_adapt.cpp:148-166). */
int adg_create(int dim, int k, int m, int nmax, adg_ctx** out);
int adg_destroy(adg_ctx* c);
const char* adg_last_error(adg_ctx* c);
/* Waits for the context's stream.  Kernels raise a device error flag instead
 * of faulting: a non-finite RK stage value (ADG_E_NONFINITE, the reference
 * drivers' blow-up check, example/02_hyperbolic_02_var_coeff_coarse_grid.cpp:389-396)
 * or a missing ancestor in the surplus transform (ADG_E_ANCESTOR).  adg_sync,
 * adg_fp_to_coe and the run drivers (every time step) report and clear it. */
int adg_sync(adg_ctx* c);

/* Grid: DGSolution/DGAdapt constructor (source/DGSolution.cpp:10-57,
 * source/DGAdapt.cpp:5-37): sparse (level sum <= level_init) or full grid. */
int adg_grid_init(adg_ctx* c, int sparse, int level_init);
/* Replace the grid by an explicit element list (levels/supports, any order). */
int adg_grid_set(adg_ctx* c, const int32_t* level, const int32_t* suppt, int64_t n);
int64_t adg_num_elem(adg_ctx* c);
/* Levels, supports (E x dim) and Hash::hash_key values, in hash order. */
int adg_grid_get(adg_ctx* c, int32_t* level, int32_t* suppt, int32_t* hash);
/* DGSolution::max_mesh_level_vec (source/DGSolution.cpp:948-959). */
int adg_max_level(adg_ctx* c, int32_t* out);
/* Hash::hash_key (source/Hash.cpp:55-114) of one element; host only. */
int32_t adg_hash_key(const int32_t* level, const int32_t* suppt, int dim);

int adg_field_set(adg_ctx* c, int field, const double* host);
int adg_field_get(adg_ctx* c, int field, double* host);

/* One 1D sweep, FastMultiplyLU::transform_1D (source/FastMultiplyLU.cpp:436-512),
 * on host arrays: out = coef * sum over related sources (vol or flx relation,
 * L/U/full filter) of in * table. */
int adg_sweep(adg_ctx* c, int table, int flx, int lu, int dim, const int32_t* size_in,
              const int32_t* size_out, double coef, const double* in, double* out);

/* Allocation-free variant of adg_sweep.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks and results as adg_sweep; the device copies of in and out are staged
 * in the library's per-GPU scratch array (2^22 doubles, held by every GPU the
 * library runs on until the process ends) instead of library-allocated
 * buffers, and the call makes no hipMalloc / hipFree.  E x (input + output
 * values per element) above 2^22 is refused (ADG_E_ARG).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other host
 * threads, other contexts) race. */
int adg_sweep_noalloc(adg_ctx* c, int table, int flx, int lu, int dim, const int32_t* size_in,
                      const int32_t* size_out, double coef, const double* in, double* out);

/* Coordinates of every element's interpolation points of the current family
 * (LagrBasis::intep_pt, source/LagrBasis.cpp:19-28; HermBasis::intep_pt,
 * source/HermBasis.cpp:9-30): x[(e*P + q)*dim + d], points in row-major
 * order.  For host flux / coefficient closures. */
int adg_lagr_points(adg_ctx* c, double* x);

/* Interpolation family of adg_eval_up, adg_flux, adg_fp_to_coe,
 * adg_rhs_lagr_vol/flx, adg_sweep tables 0-3, adg_stage and
 * adg_run_hyperbolic.  ADG_INTERP_HERMITE is the Hermite P3 flux
 * interpolation (HermInterpolation::nonlinear_Herm_2D_fast with FastHermIntp
 * and HyperbolicSameFluxHermRHS / HyperbolicDiffFluxHermRHS,
 * source/Interplation.cpp:1886-2316,3166-3316,3651-3850,4609-4621;
 * source/FastMultiplyLU.cpp:970-1069,1316-1336): degrees of freedom are the
 * value and slope of u at two points per cell; it needs m = 3, dim <= 2 and a
 * flux of u alone (ADG_FLUX_BURGERS, ADG_FLUX_SINCOS).  The adaptive
 * interpolation initialisation always uses Lagrange, as in the reference. */
enum { ADG_INTERP_LAGRANGE = 0, ADG_INTERP_HERMITE = 1 };
int adg_set_interpolation(adg_ctx* c, int family);

/* Coarse grid: the _coarse_grid variants of the reference
 * (FastMultiplyLU::transform_1D_coarse_grid, source/FastMultiplyLU.cpp:514-594;
 * FastLagrIntp::eval_up_Lagr_coarse_grid, :1367-1370;
 * HyperbolicLagrRHS::rhs_vol_scalar_coarse_grid / rhs_flx_intp_scalar_coarse_grid,
 * :1143-1217; BilinearFormAlpt::assemble_matrix_alpt_coarse_grid,
 * source/BilinearForm.cpp:89-159).  While cap >= 0, every sweep and operator
 * (adg_sweep, adg_eval_up, adg_rhs_lagr_*, adg_rhs_penalty, adg_rhs_linear,
 * adg_stage) ignores elements whose level sum exceeds cap: they are not
 * sources and their output is zero.  The point flux and the surplus transform
 * are not affected (as in the reference).  cap < 0 removes the cap. */
int adg_set_level_sum_cap(adg_ctx* c, int cap);

/* FastLagrIntp::eval_up_Lagr (source/FastMultiplyLU.cpp:1362-1365). */
int adg_eval_up(adg_ctx* c);
/* LagrInterpolation::eval_fp_Lag / eval_coe_u_Lag (source/Interplation.cpp:256-295,648-696). */
int adg_flux(adg_ctx* c, int kind, const double* params);
/* Point flux values computed by the caller (host closures of the reference's
 * nonlinear_Lagr_fast): ncomp = 1 (one flux used for every dimension) or dim
 * components of E x (m+1)^dim values each, in hash order. */
int adg_flux_set(adg_ctx* c, int ncomp, const double* fp);
/* LagrInterpolation::eval_fp_to_coe_D_Lag (source/Interplation.cpp:1222-1430). */
int adg_fp_to_coe(adg_ctx* c);
/* DGSolution::set_rhs_zero (source/DGSolution.cpp:1067-1076). */
int adg_rhs_zero(adg_ctx* c);
/* HyperbolicLagrRHS::rhs_vol_scalar / rhs_flx_intp_scalar (source/FastMultiplyLU.cpp:1125-1189). */
int adg_rhs_lagr_vol(adg_ctx* c);
int adg_rhs_lagr_flx(adg_ctx* c);
/* HyperbolicAlptRHS::rhs_flx_penalty_scalar (source/FastMultiplyLU.cpp:1269-1277):
 * Lax-Friedrichs penalty -alpha_d/2 [u][v]; equals the assembled
 * HyperbolicAlpt::assemble_matrix_flx_scalar pair (source/BilinearForm.cpp:759-769). */
int adg_rhs_penalty(adg_ctx* c, const double* alpha);
/* HyperbolicAlpt::assemble_matrix_scalar + BilinearForm::multi (source/BilinearForm.cpp:691-710),
 * applied matrix-free: rhs += A u for u_t + sum_d coef_d u_{x_d} = 0. */
int adg_rhs_linear(adg_ctx* c, const double* coef);

/* ExplicitRK::init / RK3SSP::step_stage / RK2Midpoint::step_stage /
 * ForwardEuler::step_stage (source/ODESolver.cpp:93,209-213,247-259,285-301). */
int adg_rk_init(adg_ctx* c);
int adg_rk_stage(adg_ctx* c, int scheme, int stage, double dt);
/* DGSolution::copy_ucoe_to_predict / copy_predict_to_ucoe (source/DGSolution.cpp:1151-1172). */
int adg_copy_to_predict(adg_ctx* c);
int adg_copy_predict_to_ucoe(adg_ctx* c);

/* Per-element sum of squared Alpert coefficients; DGAdapt::indicator_norm is
 * pow(sum, 0.5) (source/DGAdapt.cpp:1018-1030). */
int adg_indicator_sumsq(adg_ctx* c, double* out);
/* DGAdapt::refine / DGAdapt::coarsen (source/DGAdapt.cpp:371-404,675-688).
 * New elements start with zero coefficients. */
int adg_refine(adg_ctx* c, double eps, int64_t* n_added);
int adg_coarsen(adg_ctx* c, double eta, int64_t* n_removed);
/* Where refine and coarsen run (adg_refine, adg_coarsen and the adaptive steps
 * of adg_run_hyperbolic, and the deferring judge of the proposer modes):
 * on = 1 (default) the threshold tests, children, hole closure, hash keys,
 * the sorted element table and its fiber tables are computed on the GPU and
 * the host reads back counts only; on = 0 the host decides from
 * device-computed norms and rebuilds the element table (DGAdapt::refine /
 * coarsen and the Hash-keyed element map of DGSolution,
 * source/DGAdapt.cpp:371-404,675-688,982-1385, source/DGSolution.cpp:10-57).  Both give the same element sets, order and
 * coefficients.  A call whose norm lies within four ulps of its threshold
 * (where the reference's pow decides) runs on the host instead.
 * adg_device_adapt_stats: calls done on the device, calls handed to the host
 * and why (bits: 1 a norm at a threshold, 2 new elements beyond the table). */
int adg_set_device_adapt(adg_ctx* c, int on);
int adg_device_adapt_stats(adg_ctx* c, int64_t* device_calls, int64_t* host_fallbacks, int32_t* reasons);
/* Consistency check of the device element tables (no reference counterpart):
 * entries of the slots, instances, level sums, fiber bases, fiber levels and
 * fiber maps on the device that differ from a host build of the same grid. */
int adg_grid_check(adg_ctx* c, int64_t* mismatches);

/* DGAdaptIntp::init_adaptive_intp (source/DGAdaptIntp.cpp:35-45): adaptive
 * Lagrange interpolation of f with refinement threshold eps, then the
 * Lagrange-to-Alpert transform (FastLagrInit::eval_ucoe_Alpt_Lagr). */
int adg_init_adaptive_intp(adg_ctx* c, adg_point_fn f, void* user, double eps);

/* Built-in point functions of the benchmark configurations, usable as
 * adg_point_fn with `user` pointing to an adg_ic_params.  Coordinates are the
 * reference's unit cube [0,1]^dim.
 *   kind 0: sin(2 pi (sum_d x_d - shift))
 *   kind 1: prod_d sin(2 pi (x_d - shift))
 *   kind 2: base + ampl sin(2 pi sum_d x_d) + sum_b A_b exp(-|x - c_b|^2 / w_b^2)
 *           (periodic minimum-image distance)
 *   kind 3: inside if |offset + scale*x - center| <= radius, else outside
 *   kind 4: sum_b A_b exp(-|x - c_b|^2 / w_b^2) (periodic minimum-image distance) */
typedef struct {
  int kind;
  double shift, base, ampl;
  int nbump;
  double bump_amp[8], bump_width[8], bump_center[8][4];
  double inside, outside, radius, center[4], scale[4], offset[4];
} adg_ic_params;
double adg_ic_eval(const double* x, int dim, void* user);

/* Values at every element's interpolation points (degrees of freedom) of the
 * current family, E x (m+1)^dim in hash order (for Hermite: values and slopes,
 * as adg_lagr_points orders them) -> Alpert coefficients: the hierarchical
 * surplus transform, then the interpolation-to-Alpert partial sums
 * (FastLagrInit::eval_ucoe_Alpt_Lagr, source/FastMultiplyLU.cpp:1617-1620;
 * HermInterpolation::init_from_Herm_to_alpt, source/Interplation.cpp:5019-5030). */
int adg_init_from_points(adg_ctx* c, const double* values);

/* Separable initial function u0(x) = sum_i prod_d f_i(x_d, d): f returns the
 * factor of term ifunc in dimension d at coordinate x. */
typedef double (*adg_sep_fn)(double x, int d, int ifunc, void* user);
/* DGAdapt::init_separable_scalar_sum (source/DGAdapt.cpp:120-147): the 1D L2
 * projections of every factor (DGSolution::seperable_project,
 * source/DGSolution.cpp:150-167, with Basis::product_function,
 * source/Basis.cpp:38-59) give every element's coefficients as sums of
 * products (init_elem_separable, :169-184), added to the current ones; leaves
 * whose coefficient norm exceeds eps get all children, initialised the same
 * way, until no leaf does (refine_init_separable_scalar_sum,
 * source/DGAdapt.cpp:850-894).  One instance only. */
int adg_init_separable(adg_ctx* c, adg_sep_fn f, void* user, int nfunc, double eps);
/* Built-in separable functions, `user` pointing to an adg_ic_params:
 *   kind 0, dim 2, nfunc 2: sin(2 pi (x_0 - shift + x_1)) as
 *     sin(2 pi (x_0 - shift)) cos(2 pi x_1) + cos(2 pi (x_0 - shift)) sin(2 pi x_1)
 *   kind 1, nfunc 1: prod_d sin(2 pi (x_d - shift)) */
double adg_sep_eval(double x, int d, int ifunc, void* user);

/* Coefficient field: component d at point x. */
typedef double (*adg_coef_fn)(const double* x, int d, void* user);
/* DGAdapt::filter (source/DGAdapt.cpp:461-483): the coefficients of every
 * element with level sum >= start_level_sum are multiplied by
 * exp(-c (sum_d wave_speed[d] 2^{l_d})^2), c = damp_coef dt^2/2 pi^2. */
int adg_filter(adg_ctx* c, double damp_coef, const double* wave_speed, double dt, int start_level_sum);
/* DGAdapt::filter_local (source/DGAdapt.cpp:485-555): the same with the
 * element's local speed max over its corners of |speed(x, d)| (dim 2 or 3). */
int adg_filter_local(adg_ctx* c, double damp_coef, adg_coef_fn speed, void* user, double dt, int start_level_sum);

/* DGSolution::val (source/DGSolution.cpp:613-625, Element::val,
 * source/Element.cpp:123-147) at every point of the tensor grid x1d^dim:
 * out[i_0 * nx^(dim-1) + ... + i_{dim-1}] = u(x1d[i_0], ..., x1d[i_{dim-1}]).
 * Every element whose closed support holds the point contributes.  One
 * instance only. */
int adg_eval_tensor(adg_ctx* c, int nx, const double* x1d, double* out);
/* DGSolution::get_error_no_separable_scalar (source/DGSolution.cpp:326-335)
 * with Quad::norm_multiD (source/Quad.cpp:261-317): L1, L2 and Linf norms of
 * |u - f| by gauss_points-point Gauss-Legendre rules on the 2^L cells per
 * dimension, L the maximum active mesh level.  err[0..2] = L1, L2, Linf. */
int adg_error_no_separable(adg_ctx* c, adg_point_fn f, void* user, int gauss_points, double* err);

/* Whole-run drivers, the time loops of the reference example drivers. */
typedef struct {
  int64_t steps;          /* time steps taken */
  int64_t dof_stages;     /* sum over RK stages of active DoF */
  int64_t max_elem;       /* largest element count seen */
  double seconds;         /* wall time of the loop */
  double final_time;
  int64_t redone;         /* instance steps the threshold judge sent back (refine modes 1, 2) */
  int64_t proposed;       /* elements added by the proposer (refine modes 1, 2) */
} adg_run_stats;

/* RK3 on u_t + sum_d coef_d u_{x_d} = 0 with a fixed grid
 * (example/02_hyperbolic_01_scalar_const_coefficient.cpp:148-179). */
int adg_run_linear(adg_ctx* c, const double* coef, double dt, int64_t nsteps, adg_run_stats* st);

/* Nonlinear / variable-coefficient hyperbolic run with Lagrange flux
 * interpolation, LF penalty and (when adaptive) predict-refine-RK3-coarsen
 * steps (example/02_hyperbolic_05_burgers_adapt.cpp:223-498). */
typedef struct {
  int flux_kind;           /* ADG_FLUX_* */
  double flux_param;       /* s of ADG_FLUX_SINCOS / ADG_FLUX_VARCOEF */
  double alpha[4];         /* Lax-Friedrichs constants per dimension */
  double wave_speed[4];    /* dt = cfl / sum_d |c_d| 2^{L_d} */
  double cfl;
  double t_final;
  int adaptive;            /* 0: fixed grid */
  double eps, eta;         /* refine / coarsen thresholds */
  int64_t max_steps;       /* 0: no limit */
  double t_start;          /* time of the current solution (runs can be continued) */
  int form;                /* 0: one kernel per 1D sweep; 1: fused stages (3*dim launches per RK stage) */
} adg_problem;
int adg_run_hyperbolic(adg_ctx* c, const adg_problem* p, adg_run_stats* st);

/* Batches of independent solves.  A context may hold several instances of the
 * same discretisation (dim, k, m, nmax): their element tables are stored one
 * after another, fibers never cross instances, and every kernel serves all
 * instances in one launch.  Arrays crossing the C-ABI are ordered by instance,
 * then by hash key.  adg_run_hyperbolic advances every instance with its own
 * time step and thresholds until each reaches t_final (instances that finish
 * early leave the batch and are put back at the end). */
int adg_grid_init_batch(adg_ctx* c, int ninst, int sparse, int level_init);
/* users[i] is passed to f for the points of instance i.  Instance i refines
 * with its own threshold when adg_set_instance_thresholds set one (else eps). */
int adg_init_adaptive_intp_batch(adg_ctx* c, adg_point_fn f, void* const* users, double eps);
int adg_num_instances(adg_ctx* c);
/* Elements [first, first + count) of the union belong to instance inst. */
int adg_instance_range(adg_ctx* c, int inst, int64_t* first, int64_t* count);
/* Per-instance refine / coarsen thresholds (arrays of adg_num_instances values). */
int adg_set_instance_thresholds(adg_ctx* c, const double* eps, const double* eta);

/* Per-instance problem parameters of a batch (no reference counterpart: a
 * reference driver solves one problem).  Each array holds ninst x dim values,
 * instance-major, or is NULL for the adg_problem's own: coef_scale the c_d of
 * ADG_FLUX_VARCOEF_SCALED, alpha the Lax-Friedrichs constants of the penalty
 * (HyperbolicAlptRHS::rhs_flx_penalty_scalar, source/FastMultiplyLU.cpp:1269-1277),
 * wave_speed the speeds of the time-step rule of adg_run_hyperbolic. */
int adg_set_instance_params(adg_ctx* c, const double* coef_scale, const double* alpha, const double* wave_speed);

/* Thread mapping of the fused stage kernels (no counterpart in the
 * reference): 0 one thread per output value, 1 one thread per (element, index
 * over the non-swept dimensions), 2 one thread per (element, four such
 * indices), 3 one thread per (element, two such indices), all gathering
 * through the fiber map; 4, 5, 6 fiber tiles of 64, 128 or 256 elements: a
 * workgroup stages the inputs of whole fibers in LDS and walks the related
 * cells there, one thread per (element, index over the non-swept dimensions);
 * 7, 8, 9 the same tiles, one thread per element for all those indices.  All
 * mappings give identical results.  -1 (the default) times
 * them on the live grid the first time a kernel meets a grid-size class and
 * keeps the fastest. */
int adg_set_kernel_variant(adg_ctx* c, int variant);

/* Artificial viscosity in its matrix-free form (viscosity_option 3 of
 * example/02_hyperbolic_05_burgers_adapt.cpp:440-458).
 * adg_viscosity_elements: DGAdapt::update_viscosity_element
 * (source/DGAdapt.cpp:727-759) for every instance: childless elements at the
 * finest level whose norm exceeds the smooth-solution bound 2^((1 - k)|l|_1)
 * by more than shock_kappa decades; count = DGSolution::num_artific_visc_elem,
 * flags (may be null) 1 per viscosity element.  The set is held by element
 * key until the next call: a refine in between keeps it, elements a coarsen
 * removed are skipped.
 * adg_viscosity_flux: LargViscosInterpolation::support_gradu_Lagr_fast
 * (source/Interplation.cpp:4176-4189,579-640): for each d, coefficient * du/dx_d
 * at the Lagrange points inside a viscosity element's closed support (times the
 * number of such supports when overlap != 0), on the elements whose supports
 * meet a viscosity element's in every dimension, zero elsewhere; then the
 * surplus transform into fucoe[d].  Always on the Lagrange family; the volume
 * term (HyperbolicDiffFluxLagrRHS::rhs_vol_scalar) is adg_rhs_lagr_vol with the
 * Lagrange family selected. */
int adg_viscosity_elements(adg_ctx* c, double shock_kappa, int64_t* count, int32_t* flags);
int adg_viscosity_flux(adg_ctx* c, double coefficient, int overlap);

/* Artificial viscosity as an assembled operator A (viscosity options 1 and 2
 * of example/02_hyperbolic_05_burgers_adapt.cpp:343-420) and the IMEX43 time
 * stepper that treats it implicitly.
 * adg_viscosity_operator: ArtificialViscosity::assemble_matrix(coefficient,
 * option) (source/BilinearForm.cpp:1190-1349) over the viscosity elements of
 * the last adg_viscosity_elements call, every instance.  option 1: the
 * elements whose supports intersect the viscosity element's in every dimension
 * (ptr_vol_intp); option 2: the viscosity element and its ancestors
 * (find_all_parents).  blocks (may be null): number of coupled (row element,
 * column element) pairs.  Valid until the grid changes.
 * adg_rhs_viscosity_operator: ODESolver::add_rhs_matrix(artific_visc_operator)
 * (include/ODESolver.h:30): rhs += (A u), the product formed first.
 * adg_imex43_init: IMEX43(artific_visc_operator, dt) with the default "cg"
 * solver, then init() (source/ODESolver.cpp:883-912, include/ODESolver.h:93):
 * keeps u^n and the diagonal of I - dt alpha A.  One dt for every instance.
 * adg_imex43_stage: IMEX43::step_stage(stage) + final()
 * (source/ODESolver.cpp:994-1070).  Stages 0 and 1 are implicit only; stages
 * 2..4 take the rhs field, computed by the caller on the current u (volume,
 * flux and penalty terms), as the explicit part.  Leaves the stage value in
 * u.  Each implicit solve is Eigen's ConjugateGradient (Jacobi
 * preconditioner, tolerance 1e-12, at most 1000 iterations), per instance.
 * cg_iterations (may be null): the largest iteration count over the
 * instances; ADG_E_NOCONVERGE when a solve ends above its tolerance (the
 * reference exits). */
int adg_viscosity_operator(adg_ctx* c, double coefficient, int option, int64_t* blocks);
int adg_rhs_viscosity_operator(adg_ctx* c);
int adg_imex43_init(adg_ctx* c, double dt);
int adg_imex43_stage(adg_ctx* c, int stage, int* cg_iterations);

/* Learned refinement proposer (no reference counterpart).  The reference refines by predicting u^{n+1} with one
 * forward-Euler stage and giving every leaf whose predicted norm exceeds eps
 * its children (example/02_hyperbolic_05_burgers_adapt.cpp:246-291,
 * DGAdapt::refine, source/DGAdapt.cpp:371-404): ADG_REFINE_PREDICT, the
 * default.  ADG_REFINE_LEARNED replaces the prediction by a multilayer
 * perceptron on surplus norms of the current solution (ADG_PROPOSER_INPUTS
 * features per element, two hidden layers of 64 ReLU units, logistic output;
 * leaves with probability > threshold get their children).  ADG_REFINE_CURRENT
 * proposes the leaves whose current norm exceeds eps (the method without the
 * network).  In both, the threshold test stays as judge: after the RK step,
 * leaves above eps that lack children are refined and their instances redo
 * the step from u^n (adg_run_stats.redone counts them).
 * Features of an element: [0] log10(norm / eps) / 4, [1 + 2d], [2 + 2d] the
 * same for its same-level neighbours at j - 2 and j + 2 in dimension d
 * (periodic; levels 0 and 1 are their own neighbours, a missing one has norm
 * 0), [1 + 2 dim] sum of levels / (dim nmax), [2 + 2 dim] max level / nmax,
 * zeros after; clamped to [-4, 4].  Parameters: FP32, row-major W1[64][12],
 * b1[64], W2[64][64], b2[64], W3[64], b3.  Training (FP32, Adam, class-balanced
 * logistic loss) and inference (FP16 weights and activations, FP32
 * accumulation) run on the device. */
#define ADG_PROPOSER_INPUTS 12
#define ADG_PROPOSER_PARAMS 5057
enum { ADG_REFINE_PREDICT = 0, ADG_REFINE_LEARNED = 1, ADG_REFINE_CURRENT = 2 };
int adg_set_refine_mode(adg_ctx* c, int mode, double threshold);
/* Judge of ADG_REFINE_LEARNED / ADG_REFINE_CURRENT (no reference counterpart):
 * defer = 0 (default) redoes every instance the judge refines from u^n, so each
 * accepted step satisfies the threshold rule on the grid it was computed on;
 * defer = 1 keeps the step and gives the missed leaves their children after it
 * (zero coefficients, evolved from the next step on), which never redoes a step
 * but lets a missed leaf's children start one step late.
 * adg_proposer_stats reports the leaves refined by a deferring judge. */
int adg_set_judge_defer(adg_ctx* c, int defer);
int adg_proposer_stats(adg_ctx* c, int64_t* proposed, int64_t* redone, int64_t* deferred);
int adg_proposer_init(adg_ctx* c, uint64_t seed);
int adg_proposer_set_params(adg_ctx* c, const float* params);
int adg_proposer_get_params(adg_ctx* c, float* params);
/* gradient of the weighted logistic loss (sum of weights normalised to 1) over
 * n samples; weights may be null (all 1) */
int adg_proposer_gradient(adg_ctx* c, const float* feats, const uint8_t* labels, const float* weights, int64_t n,
                          float* grad, double* loss);
/* minibatch Adam; loss: class-balanced mean loss of the last epoch */
int adg_proposer_train(adg_ctx* c, const float* feats, const uint8_t* labels, int64_t n, int epochs, int batch,
                       float lr, uint64_t seed, double* loss);
int adg_proposer_predict(adg_ctx* c, const float* feats, int64_t n, float* prob);
/* features [E][ADG_PROPOSER_INPUTS] and leaf flags of the current grid; eps per instance */
int adg_proposer_features(adg_ctx* c, const double* eps, float* feats, int32_t* leaf);
/* Decision log: while enabled, every predict-and-refine step appends the
 * features of each leaf at the start of the step and whether the refinement
 * after the prediction gave it children.  File layout (little-endian):
 * char magic[8] = "ADGDLOG1"; int32 inputs, 0; int64 n; float32 feats[n][inputs];
 * uint8 labels[n]. */
int adg_decision_log(adg_ctx* c, int enable);
int64_t adg_decision_log_size(adg_ctx* c);
int adg_decision_log_get(adg_ctx* c, float* feats, uint8_t* labels);
int adg_decision_log_write(adg_ctx* c, const char* path);

/* Storage precision of adg_stage and of adg_run_hyperbolic's fused form (no
 * reference counterpart: the reference is FP64 throughout).  ADG_PREC_FP32_SPLIT runs the fiber transforms, point values,
 * flux and surplus transform in FP32, every 1D table and surplus weight split
 * into two FP32 parts (M = hi + lo, both applied by fused multiply-adds); the
 * right-hand side accumulation, the LF penalty, the RK combination and the
 * refine / coarsen norms stay FP64.  The per-sweep entry points are FP64. */
enum { ADG_PREC_FP64 = 0, ADG_PREC_FP32_SPLIT = 1 };
int adg_set_precision(adg_ctx* c, int precision);

/* Launch form of adg_stage and of adg_run_hyperbolic's fused form (no
 * reference counterpart).  0 (default): one launch per phase of the stage
 * (3*dim + up to 3 launches), each with the thread mapping the autotuner or
 * adg_set_kernel_variant picks.  1: one persistent cooperative launch per
 * stage: every workgroup loops over the work items of a phase -- (term, tile
 * of whole fibers) for the sweeps, points for the point flux, the surplus
 * transform and the RK update -- and the grid waits at a barrier between
 * phases; the tiles are those of kernel variant 7.  Lagrange family only (the
 * Hermite family keeps mode 0).  Mode 1 is switched off and returns an error:
 * in some runs its launch did not finish on the device (cause not found). */
int adg_set_stage_launch(adg_ctx* c, int mode);
/* Kernel launches and stages of the fused form since the context was created. */
int adg_stage_launch_count(adg_ctx* c, int64_t* launches, int64_t* stages);

/* Step graphs (no reference counterpart; the reference advances a step by
 * calling each stage in turn, source/ODESolver.cpp:285-301).  With on = 1,
 * adg_rk_step and the fused form of adg_run_hyperbolic record a whole step
 * (u^n <- u and every stage) into a HIP graph the second time the same step
 * runs on the same state -- grid, device buffers, problem, launch settings
 * and dt (or the per-instance steps' buffer) -- and replay it from then on.
 * Any change of that state, a grid change by refine or coarsen first of all,
 * makes the next step run without a graph and the one after record a new
 * graph.  Up to four recorded steps are kept.  Results are identical with and
 * without graphs; with profiling on, steps run without graphs.  on = 0 drops
 * the recorded steps. */
int adg_set_graph_step(adg_ctx* c, int on);
/* Steps recorded into graphs and steps replayed from them since creation. */
int adg_graph_step_count(adg_ctx* c, int64_t* captures, int64_t* replays);
/* One whole fused RK step of `scheme` (u^n <- u, then its stages; the same as
 * adg_rk_init followed by adg_stage for every stage) with step size dt. */
int adg_rk_step(adg_ctx* c, const adg_problem* p, int scheme, double dt);

/* Fixed thread mappings per kernel class (no reference counterpart): v[c] for
 * class c = 0..n-1 (n <= 8; classes as in adg_kernel_tuning), -1 leaves the
 * class to the autotuner.  adg_set_kernel_variant, when >= 0, overrides these.
 * All mappings give identical values. */
int adg_set_class_variants(adg_ctx* c, const int* v, int n);

/* The autotuner's decisions so far (no reference counterpart): up to cap
 * (key, variant) pairs, key = kernel class * 64 + grid-size class (log2 of the
 * element count); classes 0..3 are the FP64 eval-up phases, eval-up final,
 * right-hand-side phases and right-hand-side final, 4..7 the same in FP32.
 * *count receives the number of decisions. */
int adg_kernel_tuning(adg_ctx* c, int32_t* pairs, int cap, int* count);

/* One fused RK stage of the same operator: point values, flux, surplus
 * transform, right-hand side (volume, flux and LF penalty terms) and the stage
 * update in 3*dim launches; bit-identical to adg_eval_up + adg_flux +
 * adg_fp_to_coe + adg_rhs_zero + adg_rhs_lagr_vol + adg_rhs_lagr_flx +
 * adg_rhs_penalty + adg_rk_stage.  Uses p->flux_kind, p->flux_param, p->alpha. */
int adg_stage(adg_ctx* c, const adg_problem* p, int scheme, int stage, double dt);

/* In-memory snapshots (slots 0..3) of the element table and the Alpert
 * coefficients; restore replaces the grid and coefficients of the context. */
int adg_state_save(adg_ctx* c, int slot);
int adg_state_restore(adg_ctx* c, int slot);

/* State files: the element sets and Alpert coefficients of every instance.
 * No reference counterpart (the reference's IO writes text only,
 * source/IO.cpp:18-616, and reads nothing back); they carry golden fixtures
 * and checkpoints across machines.  Little-endian layout:
 *   char magic[8] = "ADGSTAT1"; int32 dim, k, m, nmax, ninst, 0;
 *   int64 count[ninst];
 *   per instance, per element in hash-key order: int32 level[dim], int32 suppt[dim];
 *   float64 ucoe[sum(count) * (k+1)^dim] (per element, row-major).
 * Reading checks magic, dimensions and every (level, support) and replaces the
 * context's grids and coefficients (user thresholds are cleared). */
int adg_state_write(adg_ctx* c, const char* path);
int adg_state_read(adg_ctx* c, const char* path);

/* Kernel timing: HIP events around every launch on the context's stream
 * (enable = 1 starts a fresh record).  adg_profile_get fills, per category
 * (0 sweeps: plain sweeps and the fused eval-up phases, 1 surplus transform,
 * 2 flux / fused eval-up final, 3 RK stage / fused right-hand-side final,
 * 4 other, 5 fused right-hand-side phases; arrays of 6), the summed kernel
 * milliseconds, launch count and algorithmic bytes (the bytes each kernel
 * must move: inputs read once plus outputs written once). */
int adg_profile(adg_ctx* c, int enable);
int adg_profile_get(adg_ctx* c, double* ms, int64_t* count, double* bytes);
/* Host wall time (ms, inclusive) while profiling: [0] refine/coarsen
 * decisions with their grid rebuilds, [1] grid rebuilds and uploads, [2] fused
 * plan builds, [3] whole time steps of adg_run_hyperbolic, [4] element-table
 * builds inside [1], [5] device table uploads inside [1].  ms has 6 entries. */
int adg_profile_host(adg_ctx* c, double* ms);

#ifdef __cplusplus
}
#endif
#endif
