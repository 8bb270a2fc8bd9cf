// C++ facade with the reference's class names and signatures over the C-ABI
// of libadg_hip.so (include/adg.h).  A driver written against the reference
// headers (example/02_hyperbolic_*.cpp) compiles against these headers and runs
// the hot path on the GPU: the DG solution lives in device memory owned by one
// adg_ctx per DGSolution; the other classes borrow it as the reference classes
// borrow their DGSolution pointer.
//
// Reference interfaces mirrored (include/*.h of the reference):
//   Hash.h, AllBasis.h, AlptBasis.h, LagrBasis.h, HermBasis.h, OperatorMatrix1D.h,
//   Element.h (static configuration), DGSolution.h, DGAdapt.h, DGAdaptIntp.h,
//   FastMultiplyLU.h (FastLagrIntp, HyperbolicLagrRHS, HyperbolicDiffFluxLagrRHS,
//   HyperbolicAlptRHS), Interpolation.h (LagrInterpolation, FluxFunction),
//   BilinearForm.h (BilinearForm, HyperbolicAlpt), ODESolver.h (ODESolver,
//   ForwardEuler, RK2Midpoint, RK3SSP), Timer.h, Optparser.h.
// Only the members the hyperbolic example drivers use are provided.
#pragma once
// the standard headers a driver gets through the reference's include/libs.h
#include <algorithm>
#include <array>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstring>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "adg.h"

namespace Const {
const double ROUND_OFF = 1e-13;
const double PI = 3.1415926535897932384626433832795028841971693993751058209749445923078164062;
}  // namespace Const

// Prints the C-ABI error and exits, as the reference reports errors (message + exit(1)).
void adg_facade_check(adg_ctx* c, int rc, const char* what);

class AlptBasis {
 public:
  static int PMAX;
};

class LagrBasis {
 public:
  static int PMAX;
  static int msh_case;
  static void set_interp_msh01();
};

class HermBasis {
 public:
  static int PMAX;
  static int msh_case;
  static void set_interp_msh01() {}
};

// Static configuration of the reference's Element class (include/Element.h:14-60).
class Element {
 public:
  static int PMAX_alpt, PMAX_intp, DIM, VEC_NUM;
  static std::vector<std::vector<bool>> is_intp;
};

class Interpolation {
 public:
  static int DIM, VEC_NUM;
};

class Hash {
 public:
  Hash() {}
  int hash_key(const std::vector<int>& l, const std::vector<int>& j);
};

template <class T>
class AllBasis {
 public:
  explicit AllBasis(const int max_level) : PMAX(T::PMAX), NMAX(max_level) {}
  int size() const { return (1 << NMAX) * (PMAX + 1); }
  const int PMAX;
  const int NMAX;
};

// The 1D operator tables are built inside the library context; the facade
// object only records the basis families and the boundary type.
template <class U, class V>
class OperatorMatrix1D {
 public:
  OperatorMatrix1D(const AllBasis<U>& u, const AllBasis<V>& v, const std::string& boundary_type)
      : row(u.size()), col(v.size()), boundary(boundary_type) {
    if (boundary_type != "period") {
      std::cout << "OperatorMatrix1D: only periodic boundaries are provided by the GPU library" << std::endl;
      std::exit(1);
    }
  }
  const int row, col;
  const std::string boundary;
};

// The viscosity element set of DGSolution (include/DGSolution.h:222-225): held
// by the library context; the facade keeps its size.
struct ViscosityElementSet {
  int64_t n = 0;
  // supports of the elements at the last update: xl[0], xr[0], xl[1], ... per element
  std::vector<double> supports;
  size_t size() const { return static_cast<size_t>(n); }
  bool empty() const { return n == 0; }
};

class DGSolution {
 public:
  DGSolution(const bool sparse, const int level_init, const int NMAX, AllBasis<AlptBasis>& all_bas,
             AllBasis<LagrBasis>& all_bas_Lag, AllBasis<HermBasis>& all_bas_Her, Hash& hash);
  virtual ~DGSolution();
  DGSolution(const DGSolution&) = delete;
  DGSolution& operator=(const DGSolution&) = delete;

  static int DIM;
  static int VEC_NUM;
  static std::vector<int> ind_var_vec;

  int max_mesh_level() const;
  // DGSolution::num_artific_visc_elem (source/DGSolution.cpp)
  int num_artific_visc_elem() const { return static_cast<int>(viscosity_element.n); }
  ViscosityElementSet viscosity_element;
  std::vector<int> max_mesh_level_vec() const;
  int nmax_level() const { return NMAX; }
  // DGSolution::get_error_no_separable_scalar (source/DGSolution.cpp:326-335): {L1, L2, Linf}
  std::vector<double> get_error_no_separable_scalar(std::function<double(std::vector<double>)> func,
                                                    const int gauss_points) const;
  // DGSolution::val (source/DGSolution.cpp:613-625), zero derivatives only
  std::vector<double> val(const std::vector<double>& x, const std::vector<int>& derivative) const;
  // u on the tensor grid x1d^DIM (first dimension slowest)
  std::vector<double> val_tensor(const std::vector<double>& x1d) const;
  // level / support of every element, hash-key order
  void elements(std::vector<int>& level, std::vector<int>& suppt) const;
  int size_elem() const;
  int size_basis_alpt() const;
  int get_dof() const;
  void copy_ucoe_to_predict();
  void copy_predict_to_ucoe();
  void set_rhs_zero();
  std::vector<double> get_L2_norm() const;

  // Device handle of this solution (the library context).
  adg_ctx* ctx() const { return ctx_; }
  // Alpert coefficients keyed by hash key (host copy, for measurement).
  std::unordered_map<int, std::vector<double>> coefficients_by_key() const;

  const bool sparse;
  const int level_init;
  const int NMAX;

 protected:
  adg_ctx* ctx_ = nullptr;
};

class DGAdapt : public DGSolution {
 public:
  DGAdapt(const bool sparse, const int level_init, const int NMAX, AllBasis<AlptBasis>& all_bas,
          AllBasis<LagrBasis>& all_bas_Lag, AllBasis<HermBasis>& all_bas_Her, Hash& hash, const double eps,
          const double eta, const bool is_find_ptr_alpt, const bool is_find_ptr_intp,
          const bool is_find_ptr_general = false);
  static std::vector<int> indicator_var_adapt;
  void refine();
  void coarsen();
  // DGAdapt::update_viscosity_element (source/DGAdapt.cpp:727-759)
  void update_viscosity_element(const double shock_kappa);
  // DGAdapt::init_separable_scalar_sum (source/DGAdapt.cpp:120-147)
  void init_separable_scalar_sum(std::vector<std::function<double(double, int)>> sum_func);
  // DGAdapt::filter / filter_local (source/DGAdapt.cpp:461-555)
  void filter(const double damp_coef, const std::vector<double>& wave_speed, const double dt,
              const int filter_start_level_sum);
  void filter_local(const double damp_coef, std::function<double(std::vector<double>, int)> wave_speed_func,
                    const double dt, const int filter_start_level_sum);
  const double eps;
  const double eta;
};

class DGAdaptIntp : public DGAdapt {
 public:
  DGAdaptIntp(const bool sparse, const int level_init, const int NMAX, AllBasis<AlptBasis>& all_bas,
              AllBasis<LagrBasis>& all_bas_Lag, AllBasis<HermBasis>& all_bas_Her, Hash& hash, const double eps,
              const double eta, const bool is_find_ptr_alpt, const bool is_find_ptr_intp,
              const OperatorMatrix1D<LagrBasis, LagrBasis>& matrix_Lag,
              const OperatorMatrix1D<HermBasis, HermBasis>& matrix_Her);
  void init_adaptive_intp(std::function<double(std::vector<double>, int)> func);
  double get_L2_error_split_adaptive_intp_scalar(DGSolution& exact_solu) const;
};

class FastLagrIntp {
 public:
  FastLagrIntp(DGSolution& dgsolution, const std::vector<std::vector<double>>& Lag_pt_Alpt_1D,
               const std::vector<std::vector<double>>& Lag_pt_Alpt_1D_d1);
  void eval_up_Lagr();
  void eval_up_Lagr_coarse_grid(const int mesh_nmax);
  DGSolution* const dgsolution_ptr;
};

namespace FluxFunction {
double burgers_flux_scalar(const double u);
double burgers_flux_1st_derivative_scalar(const double u);
double burgers_flux_2nd_derivative_scalar(const double u);
double sin_flux(const double u);
double sin_flux_1st_derivative(const double u);
double sin_flux_2nd_derivative(const double u);
double cos_flux(const double u);
double cos_flux_1st_derivative(const double u);
double cos_flux_2nd_derivative(const double u);
}  // namespace FluxFunction

// FastHermIntp (source/FastMultiplyLU.cpp:1316-1336): values and slopes of u
// at the Hermite points, by the library's fast transform.
class FastHermIntp {
 public:
  FastHermIntp(DGSolution& dgsolution, const std::vector<std::vector<double>>& Her_pt_Alpt_1D);
  void eval_up_Herm();
  DGSolution* const dgsolution_ptr;
};

// HermInterpolation (include/Interpolation.h:251-390), P3: the flux f(u) is
// interpolated from its values, slopes and cross slope at two points per cell.
class HermInterpolation {
 public:
  explicit HermInterpolation(DGSolution& dgsolution);
  // point tables are held by the library context; this stays empty
  std::vector<std::vector<double>> Her_pt_Alpt_1D;

  // HermInterpolation::nonlinear_Herm_2D_fast (source/Interplation.cpp:4609-4621)
  // with host closures of the reference's signatures: point values come back
  // to the host, the chain rule of eval_fp_Her_2D runs per point, the flux
  // values go back to the device.  With is_intp = {{true, false}} only the
  // first component is interpolated (the same-flux form).
  void nonlinear_Herm_2D_fast(std::function<double(std::vector<double>, int, int)> func,
                              std::function<double(std::vector<double>, int, int, int)> func_d1,
                              std::function<double(std::vector<double>, int, int, int, int)> func_d2,
                              std::vector<std::vector<bool>> is_intp, FastHermIntp& fastHerm);
  // Device flux by id (ADG_FLUX_BURGERS / ADG_FLUX_SINCOS): no host round trip.
  void nonlinear_Herm_2D_fast(int flux_id, double flux_param, std::vector<std::vector<bool>> is_intp,
                              FastHermIntp& fastHerm);
  // HermInterpolation::nonlinear_Herm_2D_PMAX5_scalar_fast (source/Interplation.cpp:5001-5017):
  // the P5 Hermite family is not provided by the GPU library; the reference
  // drivers call it only when HermBasis::PMAX == 5, which the constructor
  // already refuses.  Prints the reason and exits, as the reference does on
  // unsupported settings.
  void nonlinear_Herm_2D_PMAX5_scalar_fast(std::function<double(std::vector<double>, int, int)> func,
                                           std::function<double(std::vector<double>, int, int, int)> func_d1,
                                           std::function<double(std::vector<double>, int, int, int, int)> func_d2,
                                           std::function<double(std::vector<double>, int, int, int, int, int)> func_d3,
                                           std::function<double(std::vector<double>, int, int, int, int, int, int)> func_d4,
                                           std::vector<std::vector<bool>> is_intp, FastHermIntp& fastHerm);
  // HermInterpolation::init_from_Herm_to_alpt (source/Interplation.cpp:5019-5030):
  // func(x, i, derivative orders) at every Hermite degree of freedom (value and
  // slopes at the points of each element), the Hermite surplus transform and the
  // Hermite-to-Alpert partial sums (adg_init_from_points).
  void init_from_Herm_to_alpt(std::function<double(std::vector<double>, int, std::vector<int>)> func,
                              const OperatorMatrix1D<HermBasis, AlptBasis>& matrix);
  DGSolution* const dgsolution_ptr;
};

// HyperbolicHermRHS / HyperbolicSameFluxHermRHS / HyperbolicDiffFluxHermRHS
// (source/FastMultiplyLU.cpp:970-1069,1221-1267): volume and flux integrals of
// the Hermite-interpolated flux.  The same-flux form reads the first flux
// component in every dimension: interpolate with is_intp = {{true, false}}
// or a device flux id whose components coincide (ADG_FLUX_BURGERS).
class HyperbolicHermRHS {
 public:
  HyperbolicHermRHS(DGSolution& dgsolution, const OperatorMatrix1D<HermBasis, AlptBasis>& oper_matx_herm);
  void rhs_vol_scalar();
  void rhs_flx_intp_scalar();
  DGSolution* const dgsolution_ptr;
};

class HyperbolicSameFluxHermRHS : public HyperbolicHermRHS {
 public:
  using HyperbolicHermRHS::HyperbolicHermRHS;
};

class HyperbolicDiffFluxHermRHS : public HyperbolicHermRHS {
 public:
  using HyperbolicHermRHS::HyperbolicHermRHS;
};

// BurgersExact (include/ExactSolution.h:6-40): u(x, y, t) of
// u_t + (u^2/2)_x + (u^2/2)_y = 0 with u(x, y, 0) = a + b sin(2 pi (x + y + c)),
// before the shock (2 b t < 1 / (2 pi)).
class BurgersExact {
 public:
  BurgersExact() : a(0.), b(1.), c(0.) {}
  BurgersExact(const double a_, const double b_, const double c_) : a(a_), b(b_), c(c_) {}
  double exact_1d(const double x, const double t) const;
  double exact_2d(const double x, const double y, const double t) const;

 private:
  const double a, b, c;
  // v = sin(2 pi (x - v t)) on the characteristic through x
  double norm_1d(const double x, const double t) const;
};

class LagrInterpolation {
 public:
  explicit LagrInterpolation(DGSolution& dgsolution);
  // point tables are held by the library context; these stay empty
  std::vector<std::vector<double>> Lag_pt_Alpt_1D, Lag_pt_Alpt_1D_d1;

  // Host flux closure (the reference's signature): point values come back to
  // the host, the closure runs per point, the flux values go back to the device.
  void nonlinear_Lagr_fast(std::function<double(std::vector<double>, int, int)> func,
                           std::vector<std::vector<bool>> is_intp, FastLagrIntp& fastLagr);
  // Device flux by id (ADG_FLUX_BURGERS / ADG_FLUX_SINCOS): no host round trip.
  void nonlinear_Lagr_fast(int flux_id, double flux_param, std::vector<std::vector<bool>> is_intp,
                           FastLagrIntp& fastLagr);
  // Variable coefficient a_d(x) u with a_d = 1 + s sin(2 pi x_{(d+1) mod dim}) (device).
  void var_coeff_u_Lagr_fast(double s, std::vector<std::vector<bool>> is_intp, FastLagrIntp& fastLagr);
  // Host coefficient closure coe_func(x, d) (the reference's signature).
  void var_coeff_u_Lagr_fast(std::function<double(std::vector<double>, int)> coe_func,
                             std::vector<std::vector<bool>> is_intp, FastLagrIntp& fastLagr);
  void var_coeff_u_Lagr_fast_coarse_grid(std::function<double(std::vector<double>, int)> coe_func,
                                         std::vector<std::vector<bool>> is_intp, FastLagrIntp& fastLagr,
                                         const int mesh_nmax);
  DGSolution* const dgsolution_ptr;
};

// LargViscosInterpolation (include/Interpolation.h:220-240): the viscous flux
// nu * u_x of the artificial viscosity, interpolated on the Lagrange family.
class LargViscosInterpolation : public LagrInterpolation {
 public:
  using LagrInterpolation::LagrInterpolation;
  // LargViscosInterpolation::support_gradu_Lagr_fast (source/Interplation.cpp:4176-4189)
  void support_gradu_Lagr_fast(const ViscosityElementSet& support_element, const double coefficient,
                               std::vector<std::vector<bool>> is_intp, FastLagrIntp& fastLagr,
                               const bool overlap = false);
};

class HyperbolicLagrRHS {
 public:
  HyperbolicLagrRHS(DGSolution& dgsolution, const OperatorMatrix1D<LagrBasis, AlptBasis>& oper_matx_lagr);
  void rhs_vol_scalar();
  void rhs_flx_intp_scalar();
  void rhs_vol_scalar_coarse_grid(const int mesh_nmax);
  void rhs_flx_intp_scalar_coarse_grid(const int mesh_nmax);
  DGSolution* const dgsolution_ptr;
};

class HyperbolicDiffFluxLagrRHS : public HyperbolicLagrRHS {
 public:
  using HyperbolicLagrRHS::HyperbolicLagrRHS;
};

class HyperbolicAlptRHS {
 public:
  HyperbolicAlptRHS(DGSolution& dgsolution, const OperatorMatrix1D<AlptBasis, AlptBasis>& oper_matx_alpt);
  void rhs_flx_penalty_scalar(const std::vector<double>& lax_alpha);
  DGSolution* const dgsolution_ptr;
};

// Operators the reference assembles into Eigen sparse matrices are recorded
// here and applied matrix-free by the library.
class BilinearForm {
 public:
  explicit BilinearForm(DGSolution& dgsolution);
  virtual ~BilinearForm() {}
  // rhs += A u
  virtual void apply_add() const;
  DGSolution* const dgsolution_ptr;

 protected:
  std::vector<double> lin_coef_;            // const-coefficient advection speeds (assemble_matrix_scalar)
  std::vector<double> lft_, rgt_;           // coefficients of ulft_vjp / urgt_vjp flux terms per dimension
  bool has_lin_ = false, has_flx_ = false;
  int mesh_nmax_ = -1;                      // level-sum cap of the _coarse_grid assembly (-1: none)
  void set_cap(int mesh_nmax);
};

class HyperbolicAlpt : public BilinearForm {
 public:
  HyperbolicAlpt(DGSolution& dgsolution, const OperatorMatrix1D<AlptBasis, AlptBasis>& oper_matx_alpt);
  void assemble_matrix_scalar(const std::vector<double>& eqnCoefficient);
  void assemble_matrix_flx_scalar(const int dim, const int sign, const double coefficient);
  void assemble_matrix_scalar_coarse_grid(const std::vector<double>& eqnCoefficient, const int mesh_nmax);
  void assemble_matrix_flx_scalar_coarse_grid(const int dim, const int sign, const int mesh_nmax,
                                              const double coefficient = 1.);
};

// ArtificialViscosity (include/BilinearForm.h:376-408): the matrix of
// viscosity_coefficient * grad u . grad v over the viscosity elements' related
// elements, assembled and applied on the device.
class ArtificialViscosity : public BilinearForm {
 public:
  ArtificialViscosity(DGSolution& dgsolution, const OperatorMatrix1D<AlptBasis, AlptBasis>& oper_matx_alpt,
                      AllBasis<AlptBasis>& all_basis_alpt);
  void assemble_matrix(const double viscosity_coefficient, const int viscosity_option = 1);
  void apply_add() const override;

 private:
  bool assembled_ = false;
};

class ODESolver {
 public:
  explicit ODESolver(BilinearForm& bilinearform);
  explicit ODESolver(DGSolution& dg);
  virtual ~ODESolver() {}
  virtual void init();
  virtual void step_rk() = 0;
  virtual void step_stage(const int stage) = 0;
  virtual void final() {}
  void set_rhs_zero() {}
  void add_rhs_to_eigenvec() {}
  void add_rhs_matrix(const BilinearForm& bilinearform) { bilinearform.apply_add(); }
  BilinearForm* const dgoperator_ptr;
  DGSolution* const dgsolution_ptr;
};

class ForwardEuler : public ODESolver {
 public:
  ForwardEuler(BilinearForm& b, const double dt_) : ODESolver(b), dt(dt_) {}
  ForwardEuler(DGSolution& d, const double dt_) : ODESolver(d), dt(dt_) {}
  void step_rk() override;
  void step_stage(const int stage) override;
  const int num_stage = 1;
  const double dt;
};

class RK3SSP : public ODESolver {
 public:
  RK3SSP(BilinearForm& b, const double dt_) : ODESolver(b), dt(dt_) {}
  RK3SSP(DGSolution& d, const double dt_) : ODESolver(d), dt(dt_) {}
  void step_rk() override;
  void step_stage(const int stage) override;
  const int num_stage = 3;
  const double dt;
};

class RK2Midpoint : public ODESolver {
 public:
  RK2Midpoint(BilinearForm& b, const double dt_) : ODESolver(b), dt(dt_) {}
  RK2Midpoint(DGSolution& d, const double dt_) : ODESolver(d), dt(dt_) {}
  void step_rk() override;
  void step_stage(const int stage) override;
  const int num_stage = 2;
  const double dt;
};

// IMEX43 (include/ODESolver.h:571-635) with an ArtificialViscosity as the
// implicit part: the implicit solves (conjugate gradient, the reference's
// default "cg") run on the device.  step_stage(stage) for stages 2..4 takes
// the solution's rhs (volume, flux and add_rhs_matrix terms) as the explicit
// part.
class IMEX43 : public ODESolver {
 public:
  IMEX43(BilinearForm& bilinearForm_implicit, const double dt, const std::string linear_solver = "cg");
  void init() override;
  void step_rk() override;
  void step_stage(const int stage) override;
  const int num_stage = 5;
  const double dt;
  int last_cg_iterations = 0;
};

// Text outputs of the reference's IO class (source/IO.cpp): point values on
// the cell-centre grid of level NMAX, element centres and supports, errors.
class IO {
 public:
  explicit IO(DGSolution& dgsolution);
  void output_num(const std::string& file_name) const;
  void output_element_center(const std::string& file_name) const;
  void output_element_support(const std::string& file_name) const;
  void output_element_level_support(const std::string& file_name) const;
  void write_error_time(const double tn, const std::vector<double>& err, const int init,
                        const std::string file_name) const;
  void write_error(const int NMAX, const std::vector<double>& err, const std::string file_name) const;
  // IO::output_num_exa (source/IO.cpp:135-151, output_num_exa_1D/2D :230-245,369-388):
  // x (y) u_num u_exa on the cell centres of the finest level
  void output_num_exa(const std::string& file_name, std::function<double(std::vector<double>)> exact_solution) const;
  // IO::output_shock_support (source/IO.cpp:518-532): xl xr per dimension of
  // every viscosity element of the last DGAdapt::update_viscosity_element
  void output_shock_support(const std::string& file_name) const;
  // Binary state files (element set + Alpert coefficients; adg_state_write /
  // adg_state_read).  No reference counterpart: checkpoints and golden fixtures.
  void output_state(const std::string& file_name) const;
  void input_state(const std::string& file_name);

 private:
  DGSolution* const dgsolution_ptr;
  const int max_mesh;
  const int num_grid_point;
  const double dx;
  std::vector<double> x_1D;
};

class Timer {
 public:
  Timer() { reset(); }
  void reset() { t0_ = std::chrono::steady_clock::now(); }
  double elapsed() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }
  void time(const std::string& what) const { std::cout << what << ": " << elapsed() << " seconds" << std::endl; }

 private:
  std::chrono::steady_clock::time_point t0_;
};

// Minimal command-line parser with the reference's OptionsParser calls.
class OptionsParser {
 public:
  OptionsParser(int argc, char* argv[]) : argc_(argc), argv_(argv) {}
  void AddOption(int* v, const char* s, const char* l, const char* help) { add(v, s, l, help, 'i'); }
  void AddOption(double* v, const char* s, const char* l, const char* help) { add(v, s, l, help, 'd'); }
  void AddOption(bool* v, const char* s, const char* l, const char* help) { add(v, s, l, help, 'b'); }
  void AddOption(std::string* v, const char* s, const char* l, const char* help) { add(v, s, l, help, 's'); }
  void Parse();
  bool Good() const { return good_; }
  void PrintUsage(std::ostream& os) const;
  void PrintOptions(std::ostream& os) const;

 private:
  struct Opt { void* v; std::string s, l, help; char t; };
  void add(void* v, const char* s, const char* l, const char* help, char t) { opts_.push_back({v, s, l, help, t}); }
  int argc_;
  char** argv_;
  bool good_ = true;
  std::vector<Opt> opts_;
};
