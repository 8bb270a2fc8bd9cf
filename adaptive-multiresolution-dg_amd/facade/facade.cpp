// Implementation of the reference-named C++ facade over the C-ABI.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "adg_facade.hpp"

#include <fstream>

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

namespace {
double call_exact(const double* x, int dim, void* user) {
  auto* f = static_cast<std::function<double(std::vector<double>)>*>(user);
  return (*f)(std::vector<double>(x, x + dim));
}
}  // namespace

std::vector<double> DGSolution::get_error_no_separable_scalar(std::function<double(std::vector<double>)> func,
                                                              const int gauss_points) const {
  std::vector<double> err(3);
  adg_facade_check(ctx_, adg_error_no_separable(ctx_, call_exact, &func, gauss_points, err.data()),
                   "get_error_no_separable_scalar");
  return err;
}

std::vector<double> DGSolution::val_tensor(const std::vector<double>& x1d) const {
  size_t n = 1;
  for (int d = 0; d < DIM; ++d) n *= x1d.size();
  std::vector<double> out(n);
  adg_facade_check(ctx_, adg_eval_tensor(ctx_, static_cast<int>(x1d.size()), x1d.data(), out.data()), "DGSolution::val");
  return out;
}

std::vector<double> DGSolution::val(const std::vector<double>& x, const std::vector<int>& derivative) const {
  for (int d : derivative)
    if (d != 0) {
      std::cout << "DGSolution::val: only point values (zero derivatives) are provided" << std::endl;
      std::exit(1);
    }
  // the point is the tensor grid of one coordinate per dimension only when all are equal;
  // evaluate on the grid {x_0, ..., x_{D-1}}^D and pick the entry (x_0, ..., x_{D-1})
  const std::vector<double> grid = val_tensor(x);
  size_t idx = 0;
  for (int d = 0; d < DIM; ++d) idx = idx * x.size() + d;
  return {grid[idx]};
}

void DGSolution::elements(std::vector<int>& level, std::vector<int>& suppt) const {
  const int E = size_elem();
  level.resize(static_cast<size_t>(E) * DIM);
  suppt.resize(static_cast<size_t>(E) * DIM);
  std::vector<int32_t> key(E), lv(level.size()), sp(suppt.size());
  adg_facade_check(ctx_, adg_grid_get(ctx_, lv.data(), sp.data(), key.data()), "DGSolution::elements");
  std::copy(lv.begin(), lv.end(), level.begin());
  std::copy(sp.begin(), sp.end(), suppt.begin());
}

IO::IO(DGSolution& dgsolution)
    : dgsolution_ptr(&dgsolution), max_mesh(dgsolution.nmax_level()), num_grid_point(1 << max_mesh),
      dx(1. / num_grid_point), x_1D(num_grid_point) {
  for (int i = 0; i < num_grid_point; i++) x_1D[i] = dx * (i + 0.5);
}

void IO::output_num(const std::string& file_name) const {
  // IO::output_num_1D/2D/3D (source/IO.cpp:168-186,279-297,390-414)
  const int dim = DGSolution::DIM;
  if (dim < 1 || dim > 3) {
    std::cout << "dimension out of range in IO::output_num()" << std::endl;
    std::exit(1);
  }
  const std::vector<double> u = dgsolution_ptr->val_tensor(x_1D);
  std::ofstream output(file_name);
  const size_t n = x_1D.size();
  for (size_t i = 0; i < u.size(); ++i) {
    size_t r = i;
    std::vector<double> x(dim);
    for (int d = dim - 1; d >= 0; --d) { x[d] = x_1D[r % n]; r /= n; }
    for (int d = 0; d < dim; ++d) output << x[d] << "   ";
    output << u[i];
    if (dim == 1) output << "   ";
    output << std::endl;
  }
}

namespace {
void element_lines(const DGSolution& dg, const std::string& file_name, int what) {
  std::vector<int> level, suppt;
  dg.elements(level, suppt);
  const int D = DGSolution::DIM;
  std::ofstream output(file_name);
  for (size_t e = 0; e < level.size() / D; ++e) {
    for (int d = 0; d < D; ++d) {
      const int l = level[e * D + d], j = suppt[e * D + d];
      double xl = 0., xr = 1.;
      if (l > 1) {  // Basis::Basis support (source/Basis.cpp:21-31)
        xl = std::pow(2., -l + 1.) * (j - 1.) / 2.;
        xr = std::pow(2., -l + 1.) * (j + 1.) / 2.;
      }
      if (what == 0) output << (xl + xr) / 2. << " ";
      else if (what == 1) output << xl << " " << xr << " ";
      else output << l << " " << xl << " " << xr << " ";
    }
    output << std::endl;
  }
}
}  // namespace

void IO::output_element_center(const std::string& file_name) const { element_lines(*dgsolution_ptr, file_name, 0); }
void IO::output_element_support(const std::string& file_name) const { element_lines(*dgsolution_ptr, file_name, 1); }
void IO::output_element_level_support(const std::string& file_name) const {
  element_lines(*dgsolution_ptr, file_name, 2);
}

void IO::write_error_time(const double tn, const std::vector<double>& err, const int init,
                          const std::string file_name) const {
  std::ofstream output(file_name, init == 0 ? std::ofstream::out : (std::ofstream::out | std::ofstream::app));
  output << tn << " " << err[1] << std::endl;
}

void IO::write_error(const int NMAX, const std::vector<double>& err, const std::string file_name) const {
  std::ofstream output(file_name);
  output << NMAX << " ";
  for (auto const& x : err) output << x << " ";
  output << std::endl;
}

void IO::output_num_exa(const std::string& file_name, std::function<double(std::vector<double>)> exact_solution) const {
  const int dim = DGSolution::DIM;
  if (dim != 1 && dim != 2) {
    std::cout << "dimension out of range in IO::output_num_exa()" << std::endl;
    std::exit(1);
  }
  const std::vector<double> u = dgsolution_ptr->val_tensor(x_1D);
  std::ofstream output(file_name);
  const size_t n = x_1D.size();
  for (size_t i = 0; i < u.size(); ++i) {
    std::vector<double> x(dim);
    if (dim == 1) x[0] = x_1D[i];
    else { x[0] = x_1D[i / n]; x[1] = x_1D[i % n]; }
    for (int d = 0; d < dim; ++d) output << x[d] << "   ";
    output << u[i] << "   " << exact_solution(x) << std::endl;
  }
}

void IO::output_shock_support(const std::string& file_name) const {
  const ViscosityElementSet& v = dgsolution_ptr->viscosity_element;
  const int D = DGSolution::DIM;
  std::ofstream output(file_name);
  for (size_t e = 0; e < v.supports.size() / (2 * D); ++e) {
    for (int d = 0; d < D; ++d) output << v.supports[(e * D + d) * 2] << " " << v.supports[(e * D + d) * 2 + 1] << " ";
    output << std::endl;
  }
}

void IO::output_state(const std::string& file_name) const {
  adg_facade_check(dgsolution_ptr->ctx(), adg_state_write(dgsolution_ptr->ctx(), file_name.c_str()), "output_state");
}

void IO::input_state(const std::string& file_name) {
  adg_facade_check(dgsolution_ptr->ctx(), adg_state_read(dgsolution_ptr->ctx(), file_name.c_str()), "input_state");
}

DGAdapt::DGAdapt(const bool sparse_, const int level_init_, const int NMAX_, AllBasis<AlptBasis>& a,
                 AllBasis<LagrBasis>& l, AllBasis<HermBasis>& h, Hash& hash, const double eps_, const double eta_,
                 const bool, const bool, const bool)
    : DGSolution(sparse_, level_init_, NMAX_, a, l, h, hash), eps(eps_), eta(eta_) {}

void DGAdapt::refine() { adg_facade_check(ctx_, adg_refine(ctx_, eps, nullptr), "DGAdapt::refine"); }
void DGAdapt::coarsen() { adg_facade_check(ctx_, adg_coarsen(ctx_, eta, nullptr), "DGAdapt::coarsen"); }
void DGAdapt::update_viscosity_element(const double shock_kappa) {
  int64_t n = 0;
  std::vector<int32_t> flags(static_cast<size_t>(std::max(size_elem(), 1)));
  adg_facade_check(ctx_, adg_viscosity_elements(ctx_, shock_kappa, &n, flags.data()), "update_viscosity_element");
  viscosity_element.n = n;
  // supports of the flagged elements (Element::xl / xr, source/Element.cpp:114-120)
  std::vector<int> level, suppt;
  elements(level, suppt);
  viscosity_element.supports.clear();
  const int D = DIM;
  for (size_t e = 0; e < level.size() / D; ++e) {
    if (!flags[e]) continue;
    for (int d = 0; d < D; ++d) {
      const int l = level[e * D + d], j = suppt[e * D + d];
      double xl = 0., xr = 1.;
      if (l > 1) {
        xl = std::pow(2., -l + 1.) * (j - 1.) / 2.;
        xr = std::pow(2., -l + 1.) * (j + 1.) / 2.;
      }
      viscosity_element.supports.push_back(xl);
      viscosity_element.supports.push_back(xr);
    }
  }
}

void LargViscosInterpolation::support_gradu_Lagr_fast(const ViscosityElementSet&, const double coefficient,
                                                      std::vector<std::vector<bool>>, FastLagrIntp&,
                                                      const bool overlap) {
  // the viscosity elements are those of the last DGAdapt::update_viscosity_element on this solution
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_viscosity_flux(c, coefficient, overlap ? 1 : 0), "support_gradu_Lagr_fast");
}

namespace {
double call_separable(double x, int d, int ifunc, void* user) {
  auto* f = static_cast<std::vector<std::function<double(double, int)>>*>(user);
  return (*f)[ifunc](x, d);
}
double call_coefficient(const double* x, int d, void* user) {
  auto* f = static_cast<std::function<double(std::vector<double>, int)>*>(user);
  return (*f)(std::vector<double>(x, x + DGSolution::DIM), d);
}
// Sets the level-sum cap of the _coarse_grid variants for one call.
struct CoarseGrid {
  adg_ctx* c;
  CoarseGrid(adg_ctx* ctx, int mesh_nmax) : c(ctx) { adg_facade_check(c, adg_set_level_sum_cap(c, mesh_nmax), "coarse grid"); }
  ~CoarseGrid() { adg_set_level_sum_cap(c, -1); }
};
}  // namespace

void DGAdapt::init_separable_scalar_sum(std::vector<std::function<double(double, int)>> sum_func) {
  adg_facade_check(ctx_, adg_init_separable(ctx_, call_separable, &sum_func, static_cast<int>(sum_func.size()), eps),
                   "init_separable_scalar_sum");
}

void DGAdapt::filter(const double damp_coef, const std::vector<double>& wave_speed, const double dt,
                     const int filter_start_level_sum) {
  adg_facade_check(ctx_, adg_filter(ctx_, damp_coef, wave_speed.data(), dt, filter_start_level_sum), "DGAdapt::filter");
}

void DGAdapt::filter_local(const double damp_coef, std::function<double(std::vector<double>, int)> wave_speed_func,
                           const double dt, const int filter_start_level_sum) {
  adg_facade_check(ctx_, adg_filter_local(ctx_, damp_coef, call_coefficient, &wave_speed_func, dt, filter_start_level_sum),
                   "DGAdapt::filter_local");
}

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

void FastLagrIntp::eval_up_Lagr() {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_set_interpolation(c, ADG_INTERP_LAGRANGE), "eval_up_Lagr");
  adg_facade_check(c, adg_eval_up(c), "eval_up_Lagr");
}

void FastLagrIntp::eval_up_Lagr_coarse_grid(const int mesh_nmax) {
  CoarseGrid cg(dgsolution_ptr->ctx(), mesh_nmax);
  eval_up_Lagr();
}

double FluxFunction::burgers_flux_scalar(const double u) { return u * u * 0.5; }
double FluxFunction::burgers_flux_1st_derivative_scalar(const double u) { return u; }
double FluxFunction::burgers_flux_2nd_derivative_scalar(const double) { return 1.; }
double FluxFunction::sin_flux(const double u) { return sin(u); }
double FluxFunction::sin_flux_1st_derivative(const double u) { return cos(u); }
double FluxFunction::sin_flux_2nd_derivative(const double u) { return -sin(u); }
double FluxFunction::cos_flux(const double u) { return cos(u); }
double FluxFunction::cos_flux_1st_derivative(const double u) { return -sin(u); }
double FluxFunction::cos_flux_2nd_derivative(const double u) { return -cos(u); }

FastHermIntp::FastHermIntp(DGSolution& dg, const std::vector<std::vector<double>>&) : dgsolution_ptr(&dg) {}

void FastHermIntp::eval_up_Herm() {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_set_interpolation(c, ADG_INTERP_HERMITE), "eval_up_Herm");
  adg_facade_check(c, adg_eval_up(c), "eval_up_Herm");
}

HermInterpolation::HermInterpolation(DGSolution& dg) : dgsolution_ptr(&dg) {
  if (DGSolution::DIM != 2 || HermBasis::PMAX != 3) {
    std::cout << "HermInterpolation: the GPU library provides P3 Hermite interpolation in two dimensions" << std::endl;
    std::exit(1);
  }
}

void HermInterpolation::nonlinear_Herm_2D_fast(std::function<double(std::vector<double>, int, int)> func,
                                               std::function<double(std::vector<double>, int, int, int)> func_d1,
                                               std::function<double(std::vector<double>, int, int, int, int)> func_d2,
                                               std::vector<std::vector<bool>> is_intp, FastHermIntp& fastHerm) {
  adg_ctx* c = dgsolution_ptr->ctx();
  fastHerm.eval_up_Herm();
  const int D = 2, P = 16;
  const int64_t E = adg_num_elem(c), n = E * P;
  std::vector<double> up(n);
  adg_facade_check(c, adg_field_get(c, ADG_FIELD_UP, up.data()), "nonlinear_Herm_2D_fast");
  bool intp[2] = {true, true};
  if (!is_intp.empty())
    for (int d = 0; d < D; ++d) intp[d] = is_intp[0][d];
  const int ncomp = intp[1] ? D : 1;
  std::vector<double> fp(n * ncomp, 0.);
  std::vector<double> u(1);
  // eval_fp_Her_2D (source/Interplation.cpp:2045-2316), one scalar unknown
  for (int64_t e = 0; e < E; ++e) {
    const double* ue = &up[e * P];
    for (int q = 0; q < P; ++q) {
      const int a = q / 4, b = q % 4, k0 = a / 2, k1 = b / 2;
      for (int d = 0; d < ncomp; ++d) {
        if (!intp[d]) continue;
        double v;
        if (k0 == 0 && k1 == 0) {
          u[0] = ue[q];
          v = func(u, 0, d);
        } else if (k1 == 0) {
          u[0] = ue[(a - 2) * 4 + b];
          v = 0.;
          v += func_d1(u, 0, d, 0) * ue[q];
        } else if (k0 == 0) {
          u[0] = ue[a * 4 + b - 2];
          v = 0.;
          v += func_d1(u, 0, d, 0) * ue[q];
        } else {
          u[0] = ue[(a - 2) * 4 + b - 2];
          const double ux = ue[a * 4 + b - 2], uy = ue[(a - 2) * 4 + b];
          v = 0.;
          v += func_d2(u, 0, d, 0, 0) * ux * uy;
          v += func_d1(u, 0, d, 0) * ue[q];
        }
        fp[d * n + e * P + q] = v;
      }
    }
  }
  adg_facade_check(c, adg_flux_set(c, ncomp, fp.data()), "nonlinear_Herm_2D_fast");
  adg_facade_check(c, adg_fp_to_coe(c), "nonlinear_Herm_2D_fast");
}

void HermInterpolation::nonlinear_Herm_2D_fast(int flux_id, double flux_param, std::vector<std::vector<bool>>,
                                               FastHermIntp& fastHerm) {
  adg_ctx* c = dgsolution_ptr->ctx();
  fastHerm.eval_up_Herm();
  adg_facade_check(c, adg_flux(c, flux_id, &flux_param), "nonlinear_Herm_2D_fast");
  adg_facade_check(c, adg_fp_to_coe(c), "nonlinear_Herm_2D_fast");
}

void HermInterpolation::nonlinear_Herm_2D_PMAX5_scalar_fast(
    std::function<double(std::vector<double>, int, int)>, std::function<double(std::vector<double>, int, int, int)>,
    std::function<double(std::vector<double>, int, int, int, int)>,
    std::function<double(std::vector<double>, int, int, int, int, int)>,
    std::function<double(std::vector<double>, int, int, int, int, int, int)>, std::vector<std::vector<bool>>,
    FastHermIntp&) {
  std::cout << "HermInterpolation: P5 Hermite interpolation is not provided by the GPU library (use HermBasis::PMAX = 3)"
            << std::endl;
  std::exit(1);
}

void HermInterpolation::init_from_Herm_to_alpt(std::function<double(std::vector<double>, int, std::vector<int>)> func,
                                               const OperatorMatrix1D<HermBasis, AlptBasis>&) {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_set_interpolation(c, ADG_INTERP_HERMITE), "init_from_Herm_to_alpt");
  const int D = 2, P = 16;
  const int64_t E = adg_num_elem(c);
  std::vector<double> x(static_cast<size_t>(E) * P * D), v(static_cast<size_t>(E) * P);
  adg_facade_check(c, adg_lagr_points(c, x.data()), "init_from_Herm_to_alpt");
  // degree of freedom q = (a, b): point a % 2, derivative order a / 2 in x; b likewise in y
  std::vector<double> pt(D);
  std::vector<int> der(D);
  for (int64_t e = 0; e < E; ++e)
    for (int q = 0; q < P; ++q) {
      const int a = q / 4, b = q % 4;
      der[0] = a / 2;
      der[1] = b / 2;
      for (int d = 0; d < D; ++d) pt[d] = x[(e * P + q) * D + d];
      v[e * P + q] = func(pt, 0, der);
    }
  adg_facade_check(c, adg_init_from_points(c, v.data()), "init_from_Herm_to_alpt");
}

HyperbolicHermRHS::HyperbolicHermRHS(DGSolution& dg, const OperatorMatrix1D<HermBasis, AlptBasis>&)
    : dgsolution_ptr(&dg) {}
void HyperbolicHermRHS::rhs_vol_scalar() {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_set_interpolation(c, ADG_INTERP_HERMITE), "rhs_vol_scalar");
  adg_facade_check(c, adg_rhs_lagr_vol(c), "rhs_vol_scalar");
}
void HyperbolicHermRHS::rhs_flx_intp_scalar() {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_set_interpolation(c, ADG_INTERP_HERMITE), "rhs_flx_intp_scalar");
  adg_facade_check(c, adg_rhs_lagr_flx(c), "rhs_flx_intp_scalar");
}

// The characteristic through x carries v = sin(2 pi x0) from x0 = x - v t.
// Before the shock, g(v) = v - sin(2 pi (x - v t)) is increasing on [-1, 1]
// with g(-1) <= 0 <= g(1): Newton steps safeguarded by bisection.
double BurgersExact::norm_1d(const double x, const double t) const {
  const double pi = Const::PI;
  double lo = -1., hi = 1., v = std::sin(2. * pi * x);
  for (int it = 0; it < 200; ++it) {
    const double s = std::sin(2. * pi * (x - v * t));
    const double g = v - s;
    if (g > 0) hi = v; else lo = v;
    const double dg = 1. + 2. * pi * t * std::cos(2. * pi * (x - v * t));
    double next = v - g / dg;
    if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
    if (std::abs(next - v) < 1e-15 || hi - lo < 1e-15) return next;
    v = next;
  }
  return v;
}

double BurgersExact::exact_1d(const double x, const double t) const { return a + b * norm_1d(x - a * t + c, b * t); }

// u = a + b v(x + y - 2 a t + c, 2 b t): the 2D problem is the 1D one along x + y
double BurgersExact::exact_2d(const double x, const double y, const double t) const {
  return a + b * norm_1d(x + y - 2. * a * t + c, 2. * b * t);
}

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

namespace {
// LagrInterpolation::eval_coe_u_Lag (source/Interplation.cpp:648-696) with a host
// closure, then the surplus transform
void coe_times_u(adg_ctx* c, std::function<double(std::vector<double>, int)>& coe_func,
                 const std::vector<std::vector<bool>>& is_intp) {
  const int D = DGSolution::DIM;
  int P = 1;
  for (int d = 0; d < D; ++d) P *= LagrBasis::PMAX + 1;
  const int64_t n = adg_num_elem(c) * P;
  std::vector<double> up(n), x(n * D), fp(n * D, 0.);
  adg_facade_check(c, adg_field_get(c, ADG_FIELD_UP, up.data()), "var_coeff_u_Lagr_fast");
  adg_facade_check(c, adg_lagr_points(c, x.data()), "var_coeff_u_Lagr_fast");
  std::vector<double> pos(D);
  for (int64_t i = 0; i < n; ++i) {
    std::copy(&x[i * D], &x[i * D] + D, pos.begin());
    for (int d = 0; d < D; ++d)
      if (is_intp.empty() || is_intp[0][d]) fp[d * n + i] = coe_func(pos, d) * up[i];
  }
  adg_facade_check(c, adg_flux_set(c, D, fp.data()), "var_coeff_u_Lagr_fast");
  adg_facade_check(c, adg_fp_to_coe(c), "var_coeff_u_Lagr_fast");
}
}  // namespace

void LagrInterpolation::var_coeff_u_Lagr_fast(std::function<double(std::vector<double>, int)> coe_func,
                                              std::vector<std::vector<bool>> is_intp, FastLagrIntp& fastLagr) {
  fastLagr.eval_up_Lagr();
  coe_times_u(dgsolution_ptr->ctx(), coe_func, is_intp);
}

void LagrInterpolation::var_coeff_u_Lagr_fast_coarse_grid(std::function<double(std::vector<double>, int)> coe_func,
                                                          std::vector<std::vector<bool>> is_intp,
                                                          FastLagrIntp& fastLagr, const int mesh_nmax) {
  // source/Interplation.cpp:4215-4225: only the point evaluation is restricted
  fastLagr.eval_up_Lagr_coarse_grid(mesh_nmax);
  coe_times_u(dgsolution_ptr->ctx(), coe_func, is_intp);
}

void LagrInterpolation::var_coeff_u_Lagr_fast(double s, std::vector<std::vector<bool>>, FastLagrIntp& fastLagr) {
  adg_ctx* c = dgsolution_ptr->ctx();
  fastLagr.eval_up_Lagr();
  adg_facade_check(c, adg_flux(c, ADG_FLUX_VARCOEF, &s), "var_coeff_u_Lagr_fast");
  adg_facade_check(c, adg_fp_to_coe(c), "var_coeff_u_Lagr_fast");
}

HyperbolicLagrRHS::HyperbolicLagrRHS(DGSolution& dg, const OperatorMatrix1D<LagrBasis, AlptBasis>&)
    : dgsolution_ptr(&dg) {}
void HyperbolicLagrRHS::rhs_vol_scalar() {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_set_interpolation(c, ADG_INTERP_LAGRANGE), "rhs_vol_scalar");
  adg_facade_check(c, adg_rhs_lagr_vol(c), "rhs_vol_scalar");
}
void HyperbolicLagrRHS::rhs_flx_intp_scalar() {
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_set_interpolation(c, ADG_INTERP_LAGRANGE), "rhs_flx_intp_scalar");
  adg_facade_check(c, adg_rhs_lagr_flx(c), "rhs_flx_intp_scalar");
}
void HyperbolicLagrRHS::rhs_vol_scalar_coarse_grid(const int mesh_nmax) {
  CoarseGrid cg(dgsolution_ptr->ctx(), mesh_nmax);
  rhs_vol_scalar();
}
void HyperbolicLagrRHS::rhs_flx_intp_scalar_coarse_grid(const int mesh_nmax) {
  CoarseGrid cg(dgsolution_ptr->ctx(), mesh_nmax);
  rhs_flx_intp_scalar();
}

HyperbolicAlptRHS::HyperbolicAlptRHS(DGSolution& dg, const OperatorMatrix1D<AlptBasis, AlptBasis>&)
    : dgsolution_ptr(&dg) {}
void HyperbolicAlptRHS::rhs_flx_penalty_scalar(const std::vector<double>& lax_alpha) {
  adg_facade_check(dgsolution_ptr->ctx(), adg_rhs_penalty(dgsolution_ptr->ctx(), lax_alpha.data()), "rhs_flx_penalty_scalar");
}

// ---------------------------------------------------------------------------

BilinearForm::BilinearForm(DGSolution& dg)
    : dgsolution_ptr(&dg), lin_coef_(DGSolution::DIM, 0.), lft_(DGSolution::DIM, 0.), rgt_(DGSolution::DIM, 0.) {}

void BilinearForm::set_cap(int mesh_nmax) {
  if ((has_lin_ || has_flx_) && mesh_nmax_ != mesh_nmax) {
    std::cout << "BilinearForm: terms assembled with different coarse-grid levels are not provided" << std::endl;
    std::exit(1);
  }
  mesh_nmax_ = mesh_nmax;
}

void BilinearForm::apply_add() const {
  adg_ctx* c = dgsolution_ptr->ctx();
  CoarseGrid cg(c, mesh_nmax_);
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

void HyperbolicAlpt::assemble_matrix_scalar_coarse_grid(const std::vector<double>& c, const int mesh_nmax) {
  set_cap(mesh_nmax);
  for (int d = 0; d < DGSolution::DIM; ++d) lin_coef_[d] += c[d];
  has_lin_ = true;
}

void HyperbolicAlpt::assemble_matrix_flx_scalar_coarse_grid(const int dim, const int sign, const int mesh_nmax,
                                                            const double coefficient) {
  set_cap(mesh_nmax);
  if (sign == -1) lft_[dim] += coefficient;
  else rgt_[dim] += coefficient;
  has_flx_ = true;
}

void HyperbolicAlpt::assemble_matrix_scalar(const std::vector<double>& c) {
  set_cap(-1);
  for (int d = 0; d < DGSolution::DIM; ++d) lin_coef_[d] += c[d];
  has_lin_ = true;
}

void HyperbolicAlpt::assemble_matrix_flx_scalar(const int dim, const int sign, const double coefficient) {
  set_cap(-1);
  if (sign == -1) lft_[dim] += coefficient;
  else rgt_[dim] += coefficient;
  has_flx_ = true;
}

ArtificialViscosity::ArtificialViscosity(DGSolution& dg, const OperatorMatrix1D<AlptBasis, AlptBasis>&,
                                         AllBasis<AlptBasis>&)
    : BilinearForm(dg) {}

void ArtificialViscosity::assemble_matrix(const double viscosity_coefficient, const int viscosity_option) {
  // ArtificialViscosity::assemble_matrix (source/BilinearForm.cpp:1317-1323) over the viscosity
  // elements of the last DGAdapt::update_viscosity_element on this solution
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_viscosity_operator(c, viscosity_coefficient, viscosity_option, nullptr),
                   "ArtificialViscosity::assemble_matrix");
  assembled_ = true;
}

void ArtificialViscosity::apply_add() const {
  if (!assembled_) return;  // an empty matrix
  adg_ctx* c = dgsolution_ptr->ctx();
  adg_facade_check(c, adg_rhs_viscosity_operator(c), "ArtificialViscosity (rhs += A u)");
}

IMEX43::IMEX43(BilinearForm& implicit, const double dt_, const std::string linear_solver)
    : ODESolver(implicit), dt(dt_) {
  if (!dynamic_cast<ArtificialViscosity*>(&implicit)) {
    std::cout << "IMEX43: the GPU library provides an ArtificialViscosity implicit part only" << std::endl;
    std::exit(1);
  }
  if (linear_solver != "cg") {
    std::cout << "IMEX43: the GPU library provides the conjugate-gradient solver (\"cg\") only" << std::endl;
    std::exit(1);
  }
}

void IMEX43::init() {
  adg_facade_check(dgsolution_ptr->ctx(), adg_imex43_init(dgsolution_ptr->ctx(), dt), "IMEX43::init");
}

void IMEX43::step_stage(const int stage) {
  adg_ctx* c = dgsolution_ptr->ctx();
  const int rc = adg_imex43_stage(c, stage, &last_cg_iterations);
  if (rc == ADG_E_NOCONVERGE) {
    std::cout << "fail in IMEX43::step_stage() in stage " << stage << std::endl;
    std::exit(1);
  }
  adg_facade_check(c, rc, "IMEX43::step_stage");
}

void IMEX43::step_rk() {
  std::cout << "IMEX43::step_rk needs an assembled explicit operator, which the GPU library does not keep; "
               "call step_stage with the explicit right-hand side" << std::endl;
  std::exit(1);
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
