// Adaptive 2D Burgers with Lagrange P3 flux interpolation, written against the
// reference class API (the structure of example/02_hyperbolic_05_burgers_adapt.cpp,
// with the Lagrange classes in place of the Hermite ones).  Runs on the GPU
// through the facade.  Prints a checksum of the final coefficients.
#include "AllBasis.h"
#include "AlptBasis.h"
#include "BilinearForm.h"
#include "DGAdaptIntp.h"
#include "FastMultiplyLU.h"
#include "Hash.h"
#include "Interpolation.h"
#include "LagrBasis.h"
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
  Interpolation::DIM = DGSolution::DIM;
  Interpolation::VEC_NUM = DGSolution::VEC_NUM;

  int NMAX = 6, N_init = 4, host_flux = 0, max_steps = 0;
  double final_time = 0.1, cfl = 0.2, refine_eps = 1e-4, coarsen_eta = 1e-5;
  OptionsParser args(argc, argv);
  args.AddOption(&NMAX, "-NM", "--max-mesh-level", "Maximum mesh level");
  args.AddOption(&N_init, "-N0", "--initial-mesh-level", "Mesh level in initialization");
  args.AddOption(&final_time, "-tf", "--final-time", "Final time; start time is 0.");
  args.AddOption(&refine_eps, "-r", "--refine-epsilon", "refine parameter epsilon");
  args.AddOption(&coarsen_eta, "-c", "--coarsen-eta", "coarsen parameter eta");
  args.AddOption(&host_flux, "-hf", "--host-flux", "flux by host closure (1) or by device id (0)");
  args.AddOption(&max_steps, "-ms", "--max-steps", "stop after this many steps (0: none)");
  args.Parse();
  if (!args.Good()) { args.PrintUsage(std::cout); return 1; }

  Hash hash;
  LagrBasis::set_interp_msh01();
  AllBasis<LagrBasis> all_bas_lagr(NMAX);
  AllBasis<HermBasis> all_bas_herm(NMAX);
  AllBasis<AlptBasis> all_bas_alpt(NMAX);
  OperatorMatrix1D<AlptBasis, AlptBasis> oper_matx_alpt(all_bas_alpt, all_bas_alpt, "period");
  OperatorMatrix1D<LagrBasis, AlptBasis> oper_matx_lagr(all_bas_lagr, all_bas_alpt, "period");
  OperatorMatrix1D<LagrBasis, LagrBasis> oper_matx_lagr_lagr(all_bas_lagr, all_bas_lagr, "period");
  OperatorMatrix1D<HermBasis, HermBasis> oper_matx_herm_herm(all_bas_herm, all_bas_herm, "period");

  DGAdaptIntp dg_solu(true, N_init, NMAX, all_bas_alpt, all_bas_lagr, all_bas_herm, hash, refine_eps, coarsen_eta,
                      true, true, oper_matx_lagr_lagr, oper_matx_herm_herm);
  auto init_func = [](std::vector<double> x, int) {
    const double dx = std::min(std::abs(x[0] - 0.4), 1. - std::abs(x[0] - 0.4));
    const double dy = std::min(std::abs(x[1] - 0.6), 1. - std::abs(x[1] - 0.6));
    return 0.5 + 0.25 * std::sin(2. * Const::PI * (x[0] + x[1])) + 0.15 * std::exp(-(dx * dx + dy * dy) / (0.08 * 0.08));
  };
  dg_solu.init_adaptive_intp(init_func);

  auto func_flux = [](std::vector<double> u, int, int) -> double { return FluxFunction::burgers_flux_scalar(u[0]); };
  HyperbolicDiffFluxLagrRHS fast_rhs_lagr(dg_solu, oper_matx_lagr);
  LagrInterpolation interp_lagr(dg_solu);
  FastLagrIntp fast_lagr_intp(dg_solu, interp_lagr.Lag_pt_Alpt_1D, interp_lagr.Lag_pt_Alpt_1D_d1);
  const std::vector<std::vector<bool>> is_intp_lagr{std::vector<bool>(DIM, true)};
  auto interpolate = [&]() {
    if (host_flux) interp_lagr.nonlinear_Lagr_fast(func_flux, is_intp_lagr, fast_lagr_intp);
    else interp_lagr.nonlinear_Lagr_fast(ADG_FLUX_BURGERS, 0.0, is_intp_lagr, fast_lagr_intp);
  };
  const std::vector<double> lxf_alpha{1.2, 1.2};
  const std::vector<double> wave_speed{1., 1.};

  Timer record_time;
  double curr_time = 0.;
  int num_time_step = 0;
  while (curr_time < final_time && (max_steps <= 0 || num_time_step < max_steps)) {
    const std::vector<int> max_mesh = dg_solu.max_mesh_level_vec();
    double sum_c_dx = 0.;
    for (int d = 0; d < DIM; d++) sum_c_dx += std::abs(wave_speed[d]) * std::pow(2., max_mesh[d]);
    double dt = cfl / sum_c_dx;
    dt = std::min(dt, final_time - curr_time);
    {
      dg_solu.copy_ucoe_to_predict();
      HyperbolicAlpt linear(dg_solu, oper_matx_alpt);
      linear.assemble_matrix_flx_scalar(0, -1, lxf_alpha[0] / 2);
      linear.assemble_matrix_flx_scalar(0, 1, -lxf_alpha[0] / 2);
      linear.assemble_matrix_flx_scalar(1, -1, lxf_alpha[1] / 2);
      linear.assemble_matrix_flx_scalar(1, 1, -lxf_alpha[1] / 2);
      ForwardEuler odeSolver(linear, dt);
      odeSolver.init();
      interpolate();
      dg_solu.set_rhs_zero();
      fast_rhs_lagr.rhs_vol_scalar();
      fast_rhs_lagr.rhs_flx_intp_scalar();
      odeSolver.set_rhs_zero();
      odeSolver.add_rhs_to_eigenvec();
      odeSolver.add_rhs_matrix(linear);
      odeSolver.step_stage(0);
      odeSolver.final();
    }
    dg_solu.refine();
    dg_solu.copy_predict_to_ucoe();
    HyperbolicAlpt linear(dg_solu, oper_matx_alpt);
    linear.assemble_matrix_flx_scalar(0, -1, lxf_alpha[0] / 2);
    linear.assemble_matrix_flx_scalar(0, 1, -lxf_alpha[0] / 2);
    linear.assemble_matrix_flx_scalar(1, -1, lxf_alpha[1] / 2);
    linear.assemble_matrix_flx_scalar(1, 1, -lxf_alpha[1] / 2);
    RK3SSP odeSolver(linear, dt);
    odeSolver.init();
    for (int stage = 0; stage < odeSolver.num_stage; ++stage) {
      interpolate();
      dg_solu.set_rhs_zero();
      fast_rhs_lagr.rhs_vol_scalar();
      fast_rhs_lagr.rhs_flx_intp_scalar();
      odeSolver.set_rhs_zero();
      odeSolver.add_rhs_to_eigenvec();
      odeSolver.add_rhs_matrix(linear);
      odeSolver.step_stage(stage);
      odeSolver.final();
    }
    dg_solu.coarsen();
    curr_time += dt;
    num_time_step++;
  }
  const auto coe = dg_solu.coefficients_by_key();
  double checksum = 0.;
  for (const auto& it : coe)
    for (size_t i = 0; i < it.second.size(); ++i) checksum += (it.first % 97 + 1) * (i + 1) * it.second[i];
  std::cout.precision(17);
  std::cout << "steps " << num_time_step << " time " << curr_time << " elements " << dg_solu.size_elem()
            << " dof " << dg_solu.get_dof() << " l2sq " << dg_solu.get_L2_norm()[0] << " checksum " << checksum
            << std::endl;
  record_time.time("running time");
  return 0;
}
