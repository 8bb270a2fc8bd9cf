// 2D linear advection u_t + u_x + u_y = 0 on a sparse grid, written against the
// reference class API (the structure of
// example/02_hyperbolic_01_scalar_const_coefficient.cpp); prints the L2 error.
#include "AllBasis.h"
#include "BilinearForm.h"
#include "DGAdaptIntp.h"
#include "Hash.h"
#include "ODESolver.h"
#include "OperatorMatrix1D.h"
#include "Optparser.h"
#include "Timer.h"

int main(int argc, char* argv[]) {
  const int DIM = 2;
  AlptBasis::PMAX = 2;
  LagrBasis::PMAX = 3;
  LagrBasis::msh_case = 1;
  HermBasis::PMAX = 3;
  Element::PMAX_alpt = AlptBasis::PMAX;
  Element::PMAX_intp = LagrBasis::PMAX;
  Element::DIM = DIM;
  Element::VEC_NUM = 1;
  DGSolution::DIM = Element::DIM;
  DGSolution::VEC_NUM = Element::VEC_NUM;
  DGSolution::ind_var_vec = {0};
  DGAdapt::indicator_var_adapt = {0};
  Element::is_intp = {std::vector<bool>(DIM, true)};

  int NMAX = 5;
  double final_time = 0.1;
  const double cfl = 0.1;
  OptionsParser args(argc, argv);
  args.AddOption(&NMAX, "-N", "--max-mesh-level", "Maximum mesh level");
  args.AddOption(&final_time, "-tf", "--final-time", "Final time; start time is 0.");
  args.Parse();
  if (!args.Good()) { args.PrintUsage(std::cout); return 1; }

  Hash hash;
  LagrBasis::set_interp_msh01();
  AllBasis<AlptBasis> all_bas_alpt(NMAX);
  AllBasis<LagrBasis> all_bas_lagr(NMAX);
  AllBasis<HermBasis> all_bas_herm(NMAX);
  OperatorMatrix1D<AlptBasis, AlptBasis> oper_matx(all_bas_alpt, all_bas_alpt, "period");
  OperatorMatrix1D<HermBasis, HermBasis> oper_matx_herm_herm(all_bas_herm, all_bas_herm, "period");
  OperatorMatrix1D<LagrBasis, LagrBasis> oper_matx_lagr_lagr(all_bas_lagr, all_bas_lagr, "period");

  DGAdaptIntp dg_solu(true, NMAX, NMAX, all_bas_alpt, all_bas_lagr, all_bas_herm, hash, 1e10, -1, true, false,
                      oper_matx_lagr_lagr, oper_matx_herm_herm);
  auto init_func = [=](std::vector<double> x, int) { return std::sin(2 * Const::PI * (x[0] + x[1])); };
  dg_solu.init_adaptive_intp(init_func);

  const std::vector<double> hyperbolicConst(DIM, 1.);
  const int max_mesh = dg_solu.max_mesh_level();
  const double dx = 1. / std::pow(2., max_mesh);
  double dt = dx * cfl / DIM;
  const int total_time_step = static_cast<int>(std::ceil(final_time / dt)) + 1;
  dt = final_time / total_time_step;
  HyperbolicAlpt dg_operator(dg_solu, oper_matx);
  dg_operator.assemble_matrix_scalar(hyperbolicConst);
  RK3SSP odesolver(dg_operator, dt);
  odesolver.init();
  Timer record_time;
  for (int n = 0; n < total_time_step; n++) odesolver.step_rk();
  odesolver.final();
  record_time.time("elapsed time in evolution");

  DGAdaptIntp dg_solu_ext(true, NMAX, NMAX, all_bas_alpt, all_bas_lagr, all_bas_herm, hash, 1e-6, -1, true, false,
                          oper_matx_lagr_lagr, oper_matx_herm_herm);
  auto final_func = [=](std::vector<double> x, int) {
    return std::sin(2. * Const::PI * (x[0] + x[1] - DIM * final_time));
  };
  dg_solu_ext.init_adaptive_intp(final_func);
  std::cout.precision(10);
  std::cout << "steps " << total_time_step << " dof " << dg_solu.get_dof()
            << " L2 error at final time: " << dg_solu_ext.get_L2_error_split_adaptive_intp_scalar(dg_solu) << std::endl;
  return 0;
}
