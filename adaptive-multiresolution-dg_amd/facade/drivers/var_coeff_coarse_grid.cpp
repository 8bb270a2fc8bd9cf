// 2D variable-coefficient transport u_t + div(a u) = 0, a = (0.5 - y, x - 0.5)
// (solid-body rotation), written against the reference class API in the
// structure of example/02_hyperbolic_02_var_coeff_coarse_grid.cpp: Lagrange
// interpolation of a(x) u restricted to coarse grids of level sums N1 (stage 1)
// and N2 (stage 2) of RK2 midpoint, Lax-Friedrichs penalty assembled on the
// same coarse grids, filter_local after each step, quadrature error at the end.
#include "AllBasis.h"
#include "BilinearForm.h"
#include "DGAdaptIntp.h"
#include "FastMultiplyLU.h"
#include "Hash.h"
#include "IO.h"
#include "Interpolation.h"
#include "ODESolver.h"
#include "OperatorMatrix1D.h"
#include "Optparser.h"
#include "Timer.h"

int main(int argc, char* argv[]) {
  const int DIM = 2;
  AlptBasis::PMAX = 1;
  LagrBasis::PMAX = 2;
  LagrBasis::msh_case = 1;
  HermBasis::PMAX = 3;
  Element::PMAX_alpt = AlptBasis::PMAX;
  Element::PMAX_intp = LagrBasis::PMAX;
  Element::DIM = DIM;
  Element::VEC_NUM = 1;
  DGSolution::DIM = Element::DIM;
  DGSolution::VEC_NUM = Element::VEC_NUM;
  Interpolation::DIM = DGSolution::DIM;
  Interpolation::VEC_NUM = DGSolution::VEC_NUM;
  DGSolution::ind_var_vec = {0};
  DGAdapt::indicator_var_adapt = {0};
  Element::is_intp = {std::vector<bool>(DIM, true)};

  int NMAX = 5, N1 = 4, N2 = 5, separable = 0, gauss = 2;
  double final_time = 0.2, cfl = 0.2, filter_coef = 1.0;
  std::string out_prefix = "";
  OptionsParser args(argc, argv);
  args.AddOption(&NMAX, "-NM", "--max-mesh-level", "Maximum mesh level");
  args.AddOption(&N1, "-N1", "--max-mesh-stage-1", "Level-sum cap in stage 1");
  args.AddOption(&N2, "-N2", "--max-mesh-stage-2", "Level-sum cap in stage 2");
  args.AddOption(&final_time, "-tf", "--final-time", "Final time; start time is 0.");
  args.AddOption(&cfl, "-cfl", "--cfl-number", "CFL number");
  args.AddOption(&filter_coef, "-filter", "--filter-coefficient", "Filter coefficient");
  args.AddOption(&separable, "-sep", "--separable-init", "Separable projection (1) or adaptive interpolation (0)");
  args.AddOption(&gauss, "-g", "--num-gauss-pt-compute-error", "Gauss points per cell in the error");
  args.AddOption(&out_prefix, "-o", "--output-prefix", "Write profile and element files with this prefix");
  args.Parse();
  if (!args.Good()) { args.PrintUsage(std::cout); return 1; }

  Hash hash;
  LagrBasis::set_interp_msh01();
  AllBasis<AlptBasis> all_bas_alpt(NMAX);
  AllBasis<LagrBasis> all_bas_lagr(NMAX);
  AllBasis<HermBasis> all_bas_herm(NMAX);
  OperatorMatrix1D<AlptBasis, AlptBasis> oper_matx_alpt(all_bas_alpt, all_bas_alpt, "period");
  OperatorMatrix1D<HermBasis, HermBasis> oper_matx_herm_herm(all_bas_herm, all_bas_herm, "period");
  OperatorMatrix1D<LagrBasis, LagrBasis> oper_matx_lagr_lagr(all_bas_lagr, all_bas_lagr, "period");
  OperatorMatrix1D<LagrBasis, AlptBasis> oper_matx_lagr(all_bas_lagr, all_bas_alpt, "period");

  DGAdaptIntp dg_f(false, NMAX, NMAX, all_bas_alpt, all_bas_lagr, all_bas_herm, hash, 1e10, -1, true, false,
                   oper_matx_lagr_lagr, oper_matx_herm_herm);
  auto gauss_bump = [](double x, double y) {
    return std::exp(-(x - 0.5) * (x - 0.5) * 100. - (y - 0.6) * (y - 0.6) * 100.);
  };
  if (separable) {
    std::vector<std::function<double(double, int)>> init{
        [](double x, int d) { return d == 0 ? std::exp(-(x - 0.5) * (x - 0.5) * 100.) : std::exp(-(x - 0.6) * (x - 0.6) * 100.); }};
    dg_f.init_separable_scalar_sum(init);
  } else {
    dg_f.init_adaptive_intp([&](std::vector<double> x, int) { return gauss_bump(x[0], x[1]); });
  }

  HyperbolicLagrRHS fast_rhs_lagr(dg_f, oper_matx_lagr);
  LagrInterpolation interp(dg_f);
  FastLagrIntp fast_lagr_intp(dg_f, interp.Lag_pt_Alpt_1D, interp.Lag_pt_Alpt_1D_d1);
  const std::vector<double> max_wave_speed{0.5, 0.5};
  const std::vector<double> lxf_alpha = max_wave_speed;
  const double dx = 1. / std::pow(2., dg_f.max_mesh_level());
  double dt = dx * cfl / (max_wave_speed[0] + max_wave_speed[1]);
  const int total_time_step = static_cast<int>(std::ceil(final_time / dt)) + 1;
  dt = final_time / total_time_step;

  HyperbolicAlpt linear_stage_1(dg_f, oper_matx_alpt), linear_stage_2(dg_f, oper_matx_alpt);
  for (int d = 0; d < DIM; d++) {
    linear_stage_1.assemble_matrix_flx_scalar_coarse_grid(d, -1, N1, lxf_alpha[d] / 2);
    linear_stage_1.assemble_matrix_flx_scalar_coarse_grid(d, 1, N1, -lxf_alpha[d] / 2);
    linear_stage_2.assemble_matrix_flx_scalar_coarse_grid(d, -1, N2, lxf_alpha[d] / 2);
    linear_stage_2.assemble_matrix_flx_scalar_coarse_grid(d, 1, N2, -lxf_alpha[d] / 2);
  }
  auto coe_func = [](std::vector<double> x, int d) -> double { return d == 0 ? -x[1] + 0.5 : x[0] - 0.5; };
  const std::vector<std::vector<bool>> is_intp{std::vector<bool>(DIM, true)};
  RK2Midpoint odeSolver(linear_stage_1, dt);
  Timer record_time;
  for (int step = 0; step < total_time_step; ++step) {
    odeSolver.init();
    for (int stage = 0; stage < odeSolver.num_stage; ++stage) {
      const int cap = stage == 0 ? N1 : N2;
      interp.var_coeff_u_Lagr_fast_coarse_grid(coe_func, is_intp, fast_lagr_intp, cap);
      dg_f.set_rhs_zero();
      fast_rhs_lagr.rhs_vol_scalar_coarse_grid(cap);
      fast_rhs_lagr.rhs_flx_intp_scalar_coarse_grid(cap);
      odeSolver.set_rhs_zero();
      odeSolver.add_rhs_to_eigenvec();
      odeSolver.add_rhs_matrix(stage == 0 ? linear_stage_1 : linear_stage_2);
      odeSolver.step_stage(stage);
      odeSolver.final();
      if (stage == odeSolver.num_stage - 1) dg_f.filter_local(filter_coef, coe_func, dt, N1 + 1);
    }
  }
  // exact solution: the bump rotated by angle t about (0.5, 0.5)
  const double c = std::cos(final_time), s = std::sin(final_time);
  auto final_func = [&](std::vector<double> x) {
    const double X = x[0] - 0.5, Y = x[1] - 0.5;
    return gauss_bump(0.5 + c * X + s * Y, 0.5 - s * X + c * Y);
  };
  const std::vector<double> err = dg_f.get_error_no_separable_scalar(final_func, gauss);
  std::cout << "steps " << total_time_step << " elements " << dg_f.size_elem() << std::endl;
  std::cout << "L1, L2 and Linf error at final time: " << err[0] << ", " << err[1] << ", " << err[2] << std::endl;
  if (!out_prefix.empty()) {
    IO inout(dg_f);
    inout.output_num(out_prefix + "profile2D.txt");
    inout.output_element_center(out_prefix + "centers.txt");
    inout.write_error(NMAX, err, out_prefix + "error.txt");
  }
  return 0;
}
