/* C interface of the CPU oracle.  TEST INFRASTRUCTURE ONLY: used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker
 * and as the timed CPU baseline; never by the product library.
 *
 * Every array crossing this interface is ordered by the element hash key
 * (ascending), one element after another, each element's tensor row-major
 * (last dimension fastest) as VecMultiD stores it (include/VecMultiD.h:114-152). */
#ifndef ADG_ORACLE_H
#define ADG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;
typedef double (*or_func_t)(const double* x, int dim, void* user);

/* DGSolution/DGAdapt constructor: sparse (level sum <= level_init) or full grid. */
or_ctx* or_create(int dim, int k, int m, int nmax, int sparse, int level_init);
void or_destroy(or_ctx* c);
int64_t or_num_elem(or_ctx* c);
void or_get_grid(or_ctx* c, int32_t* level, int32_t* suppt, int32_t* hash);

/* fields: 0 ucoe_alpt, 1 rhs, 2 up_intp, 3 ucoe_tn, 4 ucoe_intp, 10+d fucoe_intp[d], 20+d fp_intp[d] */
void or_get_field(or_ctx* c, int field, double* out);
void or_set_field(or_ctx* c, int field, const double* in);

/* operator tables: 0 alpt->lagr point table, 1 lagr->alpt u_v, 2 lagr->alpt u_vx,
 * 3 lagr->alpt (ulft_vjp+urgt_vjp), 4 alpt u_v, 5 alpt u_vx, 6 alpt ulft_vjp,
 * 7 alpt urgt_vjp, 8 alpt ujp_vjp */
void or_sweep(or_ctx* c, int table, int flx, int lu, int dim, const int32_t* size_in,
              const int32_t* size_out, double coef, const double* in, double* out);

/* interpolation family of eval_up / flux / fp_to_coe / rhs_lagr_*: 0 Lagrange,
 * 1 Hermite P3 (HermInterpolation + HyperbolicHermRHS; needs m = 3, flux kinds 0
 * and 1, dim <= 2 for the flux) */
void or_set_interpolation(or_ctx* c, int family);
void or_eval_up(or_ctx* c);                                   /* FastLagrIntp::eval_up_Lagr */
/* DGAdapt::update_viscosity_element; flags (may be null): 1 per viscosity element, hash order */
int64_t or_viscosity_elements(or_ctx* c, double kappa, int32_t* flags);
/* LargViscosInterpolation::support_gradu_Lagr_fast (Lagrange family): fucoe_intp[d] */
void or_viscosity_flux(or_ctx* c, double coef, int overlap);
int64_t or_viscosity_operator(or_ctx* c, double coef, int option);
void or_viscosity_operator_coo(or_ctx* c, int64_t* rows, int64_t* cols, double* vals);
void or_rhs_viscosity_operator(or_ctx* c);
double or_product_volume_interv(or_ctx* c, int ru, int rv, int rw, int du, int dv, int nq);
void or_imex43_init(or_ctx* c, double dt);
int or_imex43_stage(or_ctx* c, int stage);
/* _coarse_grid variants: sweeps and assembled operators skip elements whose
 * level sum exceeds cap (negative: no cap). */
void or_set_level_sum_cap(or_ctx* c, int cap);
/* flux at points: kind 0 burgers u^2/2 (all dims), 1 sin/cos rotating wave scaled by p[0],
 * 2 var-coeff a_d(x)*u with a_d = 1 + p[0]*sin(2*pi*x_{(d+1)%dim}),
 * 3 var-coeff c_d a_d(x) u with c_d = p[1 + d] */
void or_flux(or_ctx* c, int kind, const double* params);
void or_fp_to_coe(or_ctx* c);                                 /* eval_fp_to_coe_D_Lag */
void or_rhs_zero(or_ctx* c);
void or_rhs_lagr_vol(or_ctx* c);                              /* HyperbolicLagrRHS::rhs_vol_scalar */
void or_rhs_lagr_flx(or_ctx* c);                              /* HyperbolicLagrRHS::rhs_flx_intp_scalar */
void or_rhs_penalty(or_ctx* c, const double* alpha);          /* HyperbolicAlptRHS::rhs_flx_penalty_scalar */
void or_rhs_linear(or_ctx* c, const double* coef);            /* matrix-free const-coefficient advection */
void or_rhs_linear_assembled(or_ctx* c, const double* coef);  /* HyperbolicAlpt assembled + SpMV */
void or_rk_init(or_ctx* c);
void or_rk_stage(or_ctx* c, int scheme, int stage, double dt); /* scheme 0 forward Euler, 3 RK3SSP */
void or_copy_to_predict(or_ctx* c);
void or_copy_predict_to_ucoe(or_ctx* c);
void or_indicator_sumsq(or_ctx* c, double* out);              /* sum of squares per element */
int64_t or_refine(or_ctx* c, double eps);
int64_t or_coarsen(or_ctx* c, double eta);
void or_max_level(or_ctx* c, int32_t* out);
void or_init_adaptive_intp(or_ctx* c, or_func_t f, void* user, double eps);
typedef double (*or_sep_t)(double x, int d, int ifunc, void* user);
typedef double (*or_coef_t)(const double* x, int d, void* user);
/* DGAdapt::init_separable_scalar_sum */
void or_init_separable(or_ctx* c, or_sep_t f, void* user, int nfunc, double eps);
/* DGSolution::val on the tensor grid x1d^dim */
void or_eval_tensor(or_ctx* c, int nx, const double* x1d, double* out);
/* DGSolution::get_error_no_separable_scalar: err = {L1, L2, Linf} */
void or_error_no_separable(or_ctx* c, or_func_t f, void* user, int points, double* err);
/* DGAdapt::filter, DGAdapt::filter_local */
void or_filter(or_ctx* c, double damp, const double* speed, double dt, int start);
void or_filter_local(or_ctx* c, double damp, or_coef_t speed, void* user, double dt, int start);
double or_l2_error_split(or_ctx* num, or_ctx* exact);
void or_set_threads(int n);

/* Inspection of the tables and structures the oracle uses. */
int64_t or_table_size(or_ctx* c, int table, int32_t* rows, int32_t* cols);
void or_table(or_ctx* c, int table, double* out);
/* related elements (vol flx=0 / flx flx=1) of element `elem` (hash order) in dimension dim, in
 * summation order; returns the count, writes up to `cap` element indices (hash order) */
int64_t or_neighbors(or_ctx* c, int64_t elem, int dim, int flx, int32_t* out, int64_t cap);
/* Basis1D values: kind 0 Alpert (degree k), 1 Lagrange (degree m), 2 Hermite (degree 3) */
double or_basis_val(or_ctx* c, int kind, int order, double x, int deriv, int sgn);
double or_lagr_point(or_ctx* c, int order);  /* point of the current interpolation family */
double or_product_volume(or_ctx* c, int kind_u, int ru, int kind_v, int cv, int du, int dv);
int32_t or_hash_key(const int32_t* level, const int32_t* suppt, int dim);
void or_surplus_tables(or_ctx* c, int32_t* anc_slot, int32_t* anc_point, double* wt);
void or_lagr_mass(or_ctx* c, double* out);
void or_gl(int n, double* x, double* w);

#ifdef __cplusplus
}
#endif
#endif
