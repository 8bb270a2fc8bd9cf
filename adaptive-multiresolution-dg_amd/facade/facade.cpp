// Implementation of the reference-named C++ facade over the C-ABI.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "adg_facade.hpp"

int AlptBasis::PMAX = 0;
int LagrBasis::PMAX = 0;
int LagrBasis::msh_case = 1;
int HermBasis::PMAX = 0;
int HermBasis::msh_case = 1;
int Element::PMAX_alpt = 0, Element::PMAX_intp = 0, Element::DIM = 0, Element::VEC_NUM = 1;
std::vector<std::vector<bool>> Element::is_intp;
int Interpolation::DIM = 0, Interpolation::VEC_NUM = 1;
int DGSolution::DIM = 0;
int DGSolution::VEC_NUM = 1;
std::vector<int> DGSolution::ind_var_vec;
std::vector<int> DGAdapt::indicator_var_adapt;

void adg_facade_check(adg_ctx* c, int rc, const char* what) {
  if (rc == ADG_OK) return;
  std::cout << what << " failed (" << rc << "): " << adg_last_error(c) << std::endl;
  std::exit(1);
}

void LagrBasis::set_interp_msh01() {
  if (msh_case != 1) {
    std::cout << "LagrBasis: only msh_case 1 is provided by the GPU library" << std::endl;
    std::exit(1);
  }
}

int Hash::hash_key(const std::vector<int>& l, const std::vector<int>& j) {
  std::vector<int32_t> lv(l.begin(), l.end()), sv(j.begin(), j.end());
  return adg_hash_key(lv.data(), sv.data(), static_cast<int>(l.size()));
}

// ---------------------------------------------------------------------------

DGSolution::DGSolution(const bool sparse_, const int level_init_, const int NMAX_, AllBasis<AlptBasis>& all_bas,
                       AllBasis<LagrBasis>&, AllBasis<HermBasis>&, Hash&)
    : sparse(sparse_), level_init(level_init_), NMAX(NMAX_) {
  if (VEC_NUM != 1) {
    std::cout << "DGSolution: the GPU library solves scalar problems (VEC_NUM = 1)" << std::endl;
    std::exit(1);
  }
  const int rc = adg_create(DIM, all_bas.PMAX, LagrBasis::PMAX, NMAX, &ctx_);
  adg_facade_check(ctx_, rc, "DGSolution");
  adg_facade_check(ctx_, adg_grid_init(ctx_, sparse ? 1 : 0, level_init), "DGSolution grid");
}

DGSolution::~DGSolution() {
  if (ctx_) adg_destroy(ctx_);
}

int DGSolution::size_elem() const { return static_cast<int>(adg_num_elem(ctx_)); }

int DGSolution::size_basis_alpt() const {
  int k1 = 1;
  for (int d = 0; d < DIM; ++d) k1 *= AlptBasis::PMAX + 1;
  return size_elem() * k1;
}

int DGSolution::get_dof() const { return size_basis_alpt() * static_cast<int>(std::max<size_t>(ind_var_vec.size(), 1)); }

std::vector<int> DGSolution::max_mesh_level_vec() const {
  std::vector<int32_t> v(DIM);
  adg_facade_check(ctx_, adg_max_level(ctx_, v.data()), "max_mesh_level_vec");
  return std::vector<int>(v.begin(), v.end());
}

int DGSolution::max_mesh_level() const {
  const std::vector<int> v = max_mesh_level_vec();
  return *std::max_element(v.begin(), v.end());
}

void DGSolution::copy_ucoe_to_predict() { adg_facade_check(ctx_, adg_copy_to_predict(ctx_), "copy_ucoe_to_predict"); }
void DGSolution::copy_predict_to_ucoe() { adg_facade_check(ctx_, adg_copy_predict_to_ucoe(ctx_), "copy_predict_to_ucoe"); }
void DGSolution::set_rhs_zero() { adg_facade_check(ctx_, adg_rhs_zero(ctx_), "set_rhs_zero"); }

std::vector<double> DGSolution::get_L2_norm() const {
  // DGSolution::get_L2_norm (source/DGSolution.cpp:970-985): sum of squared coefficients
  std::vector<double> s(size_elem());
  adg_facade_check(ctx_, adg_indicator_sumsq(ctx_, s.data()), "get_L2_norm");
  double t = 0.;
  for (double v : s) t += v;
  return {t};
}

std::unordered_map<int, std::vector<double>> DGSolution::coefficients_by_key() const {
  const int E = size_elem();
  int K = 1;
  for (int d = 0; d < DIM; ++d) K *= AlptBasis::PMAX + 1;
  std::vector<int32_t> key(E);
  std::vector<double> u(static_cast<size_t>(E) * K);
  adg_facade_check(ctx_, adg_grid_get(ctx_, nullptr, nullptr, key.data()), "grid");
  adg_facade_check(ctx_, adg_field_get(ctx_, ADG_FIELD_UCOE, u.data()), "coefficients");
  std::unordered_map<int, std::vector<double>> out;
  for (int e = 0; e < E; ++e) out[key[e]] = std::vector<double>(u.begin() + e * K, u.begin() + (e + 1) * K);
  return out;
}

DGAdapt::DGAdapt(const bool sparse_, const int level_init_, const int NMAX_, AllBasis<AlptBasis>& a,
                 AllBasis<LagrBasis>& l, AllBasis<HermBasis>& h, Hash& hash, const double eps_, const double eta_,
                 const bool, const bool, const bool)
    : DGSolution(sparse_, level_init_, NMAX_, a, l, h, hash), eps(eps_), eta(eta_) {}

void DGAdapt::refine() { adg_facade_check(ctx_, adg_refine(ctx_, eps, nullptr), "DGAdapt::refine"); }
void DGAdapt::coarsen() { adg_facade_check(ctx_, adg_coarsen(ctx_, eta, nullptr), "DGAdapt::coarsen"); }

DGAdaptIntp::DGAdaptIntp(const bool sparse_, const int level_init_, const int NMAX_, AllBasis<AlptBasis>& a,
                         AllBasis<LagrBasis>& l, AllBasis<HermBasis>& h, Hash& hash, const double eps_,
                         const double eta_, const bool fa, const bool fi, const OperatorMatrix1D<LagrBasis, LagrBasis>&,
                         const OperatorMatrix1D<HermBasis, HermBasis>&)
    : DGAdapt(sparse_, level_init_, NMAX_, a, l, h, hash, eps_, eta_, fa, fi) {}

namespace {
double call_point_function(const double* x, int dim, void* user) {
  auto* f = static_cast<std::function<double(std::vector<double>, int)>*>(user);
  return (*f)(std::vector<double>(x, x + dim), 0);
}
}  // namespace

void DGAdaptIntp::init_adaptive_intp(std::function<double(std::vector<double>, int)> func) {
  adg_facade_check(ctx_, adg_init_adaptive_intp(ctx_, call_point_function, &func, eps), "init_adaptive_intp");
}

double DGAdaptIntp::get_L2_error_split_adaptive_intp_scalar(DGSolution& exact_solu) const {
  // DGAdaptIntp::get_L2_error_split_adaptive_intp_scalar (source/DGAdaptIntp.cpp:57-115)
  const auto ex = exact_solu.coefficients_by_key();
  const auto nu = coefficients_by_key();
  double err2 = 0.0;
  for (const auto& it : ex) for (double v : it.second) err2 += v * v;
  for (const auto& it : nu) for (double v : it.second) err2 += v * v;
  for (const auto& it : nu) {
    auto f = ex.find(it.first);
    if (f == ex.end()) continue;
    for (size_t i = 0; i < it.second.size(); ++i) err2 -= 2.0 * it.second[i] * f->second[i];
  }
  if (err2 < 0.0) {
    std::cout << "negative value in DGAdaptIntp::get_L2_error_split_adaptive_intp_scalar" << std::endl;
    return err2;
  }
  return std::sqrt(std::abs(err2));
}

// ---------------------------------------------------------------------------

FastLagrIntp::FastLagrIntp(DGSolution& dg, const std::vector<std::vector<double>>&,
                           const std::vector<std::vector<double>>&)
    : dgsolution_ptr(&dg) {}

void FastLagrIntp::eval_up_Lagr() { adg_facade_check(dgsolution_ptr->ctx(), adg_eval_up(dgsolution_ptr->ctx()), "eval_up_Lagr"); }

double FluxFunction::burgers_flux_scalar(const double u) { return u * u * 0.5; }
double FluxFunction::sin_flux(const double u) { return sin(u); }
double FluxFunction::cos_flux(const double u) { return cos(u); }

LagrInterpolation::LagrInterpolation(DGSolution& dg) : dgsolution_ptr(&dg) {}

void LagrInterpolation::nonlinear_Lagr_fast(std::function<double(std::vector<double>, int, int)> func,
                                            std::vector<std::vector<bool>> is_intp, FastLagrIntp& fastLagr) {
  adg_ctx* c = dgsolution_ptr->ctx();
  fastLagr.eval_up_Lagr();
  const int D = DGSolution::DIM;
  int P = 1;
  for (int d = 0; d < D; ++d) P *= LagrBasis::PMAX + 1;
  const int64_t n = adg_num_elem(c) * P;
  std::vector<double> up(n), fp(n * D);
  adg_facade_check(c, adg_field_get(c, ADG_FIELD_UP, up.data()), "nonlinear_Lagr_fast");
  std::vector<double> val(1);
  for (int64_t i = 0; i < n; ++i) {
    val[0] = up[i];
    for (int d = 0; d < D; ++d)
      if (is_intp.empty() || is_intp[0][d]) fp[d * n + i] = func(val, 0, d);
  }
  adg_facade_check(c, adg_flux_set(c, D, fp.data()), "nonlinear_Lagr_fast");
  adg_facade_check(c, adg_fp_to_coe(c), "nonlinear_Lagr_fast");
}

void LagrInterpolation::nonlinear_Lagr_fast(int flux_id, double flux_param, std::vector<std::vector<bool>>,
                                            FastLagrIntp& fastLagr) {
  adg_ctx* c = dgsolution_ptr->ctx();
  fastLagr.eval_up_Lagr();
  adg_facade_check(c, adg_flux(c, flux_id, &flux_param), "nonlinear_Lagr_fast");
  adg_facade_check(c, adg_fp_to_coe(c), "nonlinear_Lagr_fast");
}

void LagrInterpolation::var_coeff_u_Lagr_fast(double s, std::vector<std::vector<bool>>, FastLagrIntp& fastLagr) {
  adg_ctx* c = dgsolution_ptr->ctx();
  fastLagr.eval_up_Lagr();
  adg_facade_check(c, adg_flux(c, ADG_FLUX_VARCOEF, &s), "var_coeff_u_Lagr_fast");
  adg_facade_check(c, adg_fp_to_coe(c), "var_coeff_u_Lagr_fast");
}

HyperbolicLagrRHS::HyperbolicLagrRHS(DGSolution& dg, const OperatorMatrix1D<LagrBasis, AlptBasis>&)
    : dgsolution_ptr(&dg) {}
void HyperbolicLagrRHS::rhs_vol_scalar() { adg_facade_check(dgsolution_ptr->ctx(), adg_rhs_lagr_vol(dgsolution_ptr->ctx()), "rhs_vol_scalar"); }
void HyperbolicLagrRHS::rhs_flx_intp_scalar() { adg_facade_check(dgsolution_ptr->ctx(), adg_rhs_lagr_flx(dgsolution_ptr->ctx()), "rhs_flx_intp_scalar"); }

HyperbolicAlptRHS::HyperbolicAlptRHS(DGSolution& dg, const OperatorMatrix1D<AlptBasis, AlptBasis>&)
    : dgsolution_ptr(&dg) {}
void HyperbolicAlptRHS::rhs_flx_penalty_scalar(const std::vector<double>& lax_alpha) {
  adg_facade_check(dgsolution_ptr->ctx(), adg_rhs_penalty(dgsolution_ptr->ctx(), lax_alpha.data()), "rhs_flx_penalty_scalar");
}

// ---------------------------------------------------------------------------

BilinearForm::BilinearForm(DGSolution& dg)
    : dgsolution_ptr(&dg), lin_coef_(DGSolution::DIM, 0.), lft_(DGSolution::DIM, 0.), rgt_(DGSolution::DIM, 0.) {}

void BilinearForm::apply_add() const {
  adg_ctx* c = dgsolution_ptr->ctx();
  if (has_lin_) adg_facade_check(c, adg_rhs_linear(c, lin_coef_.data()), "BilinearForm (advection)");
  if (has_flx_) {
    // coef_l ulft_vjp + coef_r urgt_vjp with coef_l = -coef_r is coef_r ujp_vjp, the LF penalty
    std::vector<double> alpha(DGSolution::DIM, 0.);
    for (int d = 0; d < DGSolution::DIM; ++d) {
      if (lft_[d] != -rgt_[d]) {
        std::cout << "BilinearForm: only the Lax-Friedrichs pair assemble_matrix_flx_scalar(d,-1,a/2) + (d,1,-a/2) "
                     "is provided by the GPU library" << std::endl;
        std::exit(1);
      }
      alpha[d] = -2. * rgt_[d];
    }
    adg_facade_check(c, adg_rhs_penalty(c, alpha.data()), "BilinearForm (penalty)");
  }
}

HyperbolicAlpt::HyperbolicAlpt(DGSolution& dg, const OperatorMatrix1D<AlptBasis, AlptBasis>&) : BilinearForm(dg) {}

void HyperbolicAlpt::assemble_matrix_scalar(const std::vector<double>& c) {
  for (int d = 0; d < DGSolution::DIM; ++d) lin_coef_[d] += c[d];
  has_lin_ = true;
}

void HyperbolicAlpt::assemble_matrix_flx_scalar(const int dim, const int sign, const double coefficient) {
  if (sign == -1) lft_[dim] += coefficient;
  else rgt_[dim] += coefficient;
  has_flx_ = true;
}

ODESolver::ODESolver(BilinearForm& b) : dgoperator_ptr(&b), dgsolution_ptr(b.dgsolution_ptr) {}
ODESolver::ODESolver(DGSolution& dg) : dgoperator_ptr(nullptr), dgsolution_ptr(&dg) {}
void ODESolver::init() { adg_facade_check(dgsolution_ptr->ctx(), adg_rk_init(dgsolution_ptr->ctx()), "ODESolver::init"); }

void ForwardEuler::step_stage(const int stage) {
  adg_facade_check(dgsolution_ptr->ctx(), adg_rk_stage(dgsolution_ptr->ctx(), ADG_RK_EULER, stage, dt), "ForwardEuler");
}

void ForwardEuler::step_rk() {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_rk_init(c), "ForwardEuler::step_rk");
  adg_facade_check(c, adg_rhs_zero(c), "ForwardEuler::step_rk");
  dgoperator_ptr->apply_add();
  step_stage(0);
}

void RK3SSP::step_stage(const int stage) {
  adg_facade_check(dgsolution_ptr->ctx(), adg_rk_stage(dgsolution_ptr->ctx(), ADG_RK3, stage, dt), "RK3SSP");
}

void RK3SSP::step_rk() {
  // RK3SSP::step_rk (source/ODESolver.cpp:261-271): three stages of rhs = A u
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_rk_init(c), "RK3SSP::step_rk");
  for (int s = 0; s < 3; ++s) {
    adg_facade_check(c, adg_rhs_zero(c), "RK3SSP::step_rk");
    dgoperator_ptr->apply_add();
    step_stage(s);
  }
}

void RK2Midpoint::step_stage(const int stage) {
  adg_facade_check(dgsolution_ptr->ctx(), adg_rk_stage(dgsolution_ptr->ctx(), ADG_RK2, stage, dt), "RK2Midpoint");
}

void RK2Midpoint::step_rk() {
  // RK2Midpoint::step_rk (source/ODESolver.cpp:238-245): u1 = u_n + dt/2 A u_n, u = u_n + dt A u1
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_rk_init(c), "RK2Midpoint::step_rk");
  for (int s = 0; s < 2; ++s) {
    adg_facade_check(c, adg_rhs_zero(c), "RK2Midpoint::step_rk");
    dgoperator_ptr->apply_add();
    step_stage(s);
  }
}

// ---------------------------------------------------------------------------

void OptionsParser::Parse() {
  for (int i = 1; i < argc_; ++i) {
    const std::string a = argv_[i];
    bool found = false;
    for (Opt& o : opts_) {
      if (a != o.s && a != o.l) continue;
      found = true;
      if (o.t == 'b') { *static_cast<bool*>(o.v) = true; break; }
      if (i + 1 >= argc_) { good_ = false; return; }
      const char* val = argv_[++i];
      if (o.t == 'i') *static_cast<int*>(o.v) = std::atoi(val);
      else if (o.t == 'd') *static_cast<double*>(o.v) = std::atof(val);
      else *static_cast<std::string*>(o.v) = val;
      break;
    }
    if (!found) { good_ = false; return; }
  }
}

void OptionsParser::PrintUsage(std::ostream& os) const {
  os << "Usage: " << argv_[0] << " [options]\n";
  for (const Opt& o : opts_) os << "  " << o.s << ", " << o.l << " : " << o.help << "\n";
}

void OptionsParser::PrintOptions(std::ostream& os) const {
  os << "Options used:\n";
  for (const Opt& o : opts_) {
    os << "  " << o.l << " ";
    if (o.t == 'i') os << *static_cast<int*>(o.v);
    else if (o.t == 'd') os << *static_cast<double*>(o.v);
    else if (o.t == 'b') os << (*static_cast<bool*>(o.v) ? "true" : "false");
    else os << *static_cast<std::string*>(o.v);
    os << "\n";
  }
}
