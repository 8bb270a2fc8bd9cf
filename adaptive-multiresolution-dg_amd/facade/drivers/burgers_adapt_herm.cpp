// Adaptive 2D Burgers with Hermite P3 flux interpolation, written against the
// reference class API: the default path of
// example/02_hyperbolic_05_burgers_adapt.cpp (separable initial value
// sin(2 pi (x + y)), same-flux Hermite interpolation with is_intp {true, false},
// forward-Euler prediction, refine, SSP RK3, coarsen), with its artificial
// viscosity options: 1 and 2 (assembled viscosity operator, IMEX43 or explicit
// SSP RK3) and 3 (viscous flux interpolated on the Lagrange family, explicit
// SSP RK3), then the error against BurgersExact.  Runs on the GPU through the
// facade.
#include "AllBasis.h"
#include "AlptBasis.h"
#include "BilinearForm.h"
#include "DGAdapt.h"
#include "ExactSolution.h"
#include "FastMultiplyLU.h"
#include "Hash.h"
#include "HermBasis.h"
#include "IO.h"
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
  HermBasis::msh_case = 1;
  Element::PMAX_alpt = AlptBasis::PMAX;
  Element::PMAX_intp = HermBasis::PMAX;
  Element::DIM = DIM;
  Element::VEC_NUM = 1;
  DGSolution::DIM = Element::DIM;
  DGSolution::VEC_NUM = Element::VEC_NUM;
  DGSolution::ind_var_vec = {0};
  DGAdapt::indicator_var_adapt = {0};
  Element::is_intp = {std::vector<bool>(DIM, true)};
  Interpolation::DIM = DGSolution::DIM;
  Interpolation::VEC_NUM = DGSolution::VEC_NUM;

  int NMAX = 3, N_init = 3, is_init_sparse = 1, host_flux = 0, max_steps = 0, num_gauss_pt_compute_error = 1;
  double final_time = 0.03, cfl = 0.05, cfl_hyper = 0.2, refine_eps = 1e-2, coarsen_eta = -1.;
  int viscosity_option = 0, support_overlap = 0, is_imex_option = 1;
  double shock_kappa = 0.0, artificial_viscosity_nu0 = 1.;
  OptionsParser args(argc, argv);
  args.AddOption(&NMAX, "-NM", "--max-mesh-level", "Maximum mesh level");
  args.AddOption(&N_init, "-N0", "--initial-mesh-level", "Mesh level in initialization");
  args.AddOption(&is_init_sparse, "-s", "--sparse-grid-initial", "Use full grid (0) or sparse grid (1) in initialization");
  args.AddOption(&final_time, "-tf", "--final-time", "Final time; start time is 0.");
  args.AddOption(&cfl, "-cfl", "--cfl-number", "CFL number when there are viscosity elements");
  args.AddOption(&viscosity_option, "-v", "--artificial-viscosity-option", "artificial viscosity option (0 or 3)");
  args.AddOption(&artificial_viscosity_nu0, "-nu", "--artificial-viscosity-value", "artificial viscosity nu0");
  args.AddOption(&shock_kappa, "-k", "--shock-kappa", "shock detector constant kappa");
  args.AddOption(&is_imex_option, "-imex", "--implicit-explicit", "Use IMEX (1) or explicit SSP RK (0)");
  args.AddOption(&support_overlap, "-sp", "--support-viscosity-overlap", "consider overlap of support (1) or not (0)");
  args.AddOption(&refine_eps, "-r", "--refine-epsilon", "refine parameter epsilon");
  args.AddOption(&coarsen_eta, "-c", "--coarsen-eta", "coarsen parameter eta (negative: refine_eps / 10)");
  args.AddOption(&num_gauss_pt_compute_error, "-g", "--num-gauss-pt-compute-error", "number of gauss points in computing error");
  args.AddOption(&host_flux, "-hf", "--host-flux", "flux by host closures (1) or by device id (0)");
  args.AddOption(&max_steps, "-ms", "--max-steps", "stop after this many steps (0: none)");
  std::string state_file;
  args.AddOption(&state_file, "-st", "--state-file", "write the final state to this file (empty: none)");
  args.Parse();
  if (!args.Good()) { args.PrintUsage(std::cout); return 1; }
  if (N_init > NMAX) { std::cout << "Mesh level in initialization should not be larger than Maximum mesh level" << std::endl; return 1; }
  if (coarsen_eta < 0) coarsen_eta = refine_eps / 10.;
  if (viscosity_option < 0 || viscosity_option > 3) { std::cout << "unknown artificial viscosity option" << std::endl; return 1; }
  if (is_imex_option == 1 && viscosity_option == 3) {
    std::cout << "In artificial viscosity option 3, we do not assemble matrix thus IMEX could not be applied" << std::endl;
    return 1;
  }

  Hash hash;
  LagrBasis::set_interp_msh01();
  HermBasis::set_interp_msh01();
  AllBasis<AlptBasis> all_bas_alpt(NMAX);
  AllBasis<LagrBasis> all_bas_lagr(NMAX);
  AllBasis<HermBasis> all_bas_herm(NMAX);
  OperatorMatrix1D<AlptBasis, AlptBasis> oper_matx_alpt(all_bas_alpt, all_bas_alpt, "period");
  OperatorMatrix1D<HermBasis, AlptBasis> oper_matx_herm(all_bas_herm, all_bas_alpt, "period");
  OperatorMatrix1D<LagrBasis, AlptBasis> oper_matx_lagr(all_bas_lagr, all_bas_alpt, "period");

  DGAdapt dg_solu(is_init_sparse == 1, N_init, NMAX, all_bas_alpt, all_bas_lagr, all_bas_herm, hash, refine_eps,
                  coarsen_eta, true, true);
  // sin(2 pi (x + y)) = sin(2 pi x) cos(2 pi y) + cos(2 pi x) sin(2 pi y)
  auto init_func_1 = [](double x, int d) { return (d == 0) ? std::sin(2. * Const::PI * x) : std::cos(2. * Const::PI * x); };
  auto init_func_2 = [](double x, int d) { return (d == 0) ? std::cos(2. * Const::PI * x) : std::sin(2. * Const::PI * x); };
  dg_solu.init_separable_scalar_sum({init_func_1, init_func_2});
  BurgersExact burgers(0., 1., 0.);

  auto func_flux = [](std::vector<double> u, int, int) -> double { return FluxFunction::burgers_flux_scalar(u[0]); };
  auto func_flux_d1 = [](std::vector<double> u, int, int, int) -> double {
    return FluxFunction::burgers_flux_1st_derivative_scalar(u[0]);
  };
  auto func_flux_d2 = [](std::vector<double> u, int, int, int, int) -> double {
    return FluxFunction::burgers_flux_2nd_derivative_scalar(u[0]);
  };
  HyperbolicSameFluxHermRHS fast_rhs_herm(dg_solu, oper_matx_herm);
  HermInterpolation interp_herm(dg_solu);
  const std::vector<std::vector<bool>> is_intp_herm{std::vector<bool>{true, false}};
  FastHermIntp fast_herm_intp(dg_solu, interp_herm.Her_pt_Alpt_1D);
  // artificial viscosity: nu * u_x interpolated on the Lagrange family
  HyperbolicDiffFluxLagrRHS fast_rhs_lagr(dg_solu, oper_matx_lagr);
  LargViscosInterpolation interp_lagr(dg_solu);
  FastLagrIntp fast_lagr_intp(dg_solu, interp_lagr.Lag_pt_Alpt_1D, interp_lagr.Lag_pt_Alpt_1D_d1);
  const std::vector<std::vector<bool>> is_intp_lagr{std::vector<bool>(DIM, true)};
  int total_visc_elem = 0, total_cg_iterations = 0;
  auto interpolate = [&]() {
    if (host_flux) interp_herm.nonlinear_Herm_2D_fast(func_flux, func_flux_d1, func_flux_d2, is_intp_herm, fast_herm_intp);
    else interp_herm.nonlinear_Herm_2D_fast(ADG_FLUX_BURGERS, 0.0, is_intp_herm, fast_herm_intp);
  };
  const std::vector<double> lxf_alpha{1.2, 1.2};
  const std::vector<double> wave_speed{1., 1.};
  auto lax_friedrichs = [&](HyperbolicAlpt& linear) {
    linear.assemble_matrix_flx_scalar(0, -1, lxf_alpha[0] / 2);
    linear.assemble_matrix_flx_scalar(0, 1, -lxf_alpha[0] / 2);
    linear.assemble_matrix_flx_scalar(1, -1, lxf_alpha[1] / 2);
    linear.assemble_matrix_flx_scalar(1, 1, -lxf_alpha[1] / 2);
  };

  Timer record_time;
  double curr_time = 0.;
  int num_time_step = 0;
  while (curr_time < final_time && (max_steps <= 0 || num_time_step < max_steps)) {
    int num_visc_elem = 0;
    if (viscosity_option != 0) {
      dg_solu.update_viscosity_element(shock_kappa);
      num_visc_elem = dg_solu.num_artific_visc_elem();
      total_visc_elem += num_visc_elem;
    }
    const std::vector<int> max_mesh = dg_solu.max_mesh_level_vec();
    double sum_c_dx = 0.;
    for (int d = 0; d < DIM; d++) sum_c_dx += std::abs(wave_speed[d]) * std::pow(2., max_mesh[d]);
    double dt = cfl / sum_c_dx;
    if (num_visc_elem == 0) dt = cfl_hyper / sum_c_dx;
    dt = std::min(dt, final_time - curr_time);
    {
      dg_solu.copy_ucoe_to_predict();
      HyperbolicAlpt linear(dg_solu, oper_matx_alpt);
      lax_friedrichs(linear);
      ForwardEuler odeSolver(linear, dt);
      odeSolver.init();
      interpolate();
      dg_solu.set_rhs_zero();
      fast_rhs_herm.rhs_vol_scalar();
      fast_rhs_herm.rhs_flx_intp_scalar();
      odeSolver.set_rhs_zero();
      odeSolver.add_rhs_to_eigenvec();
      odeSolver.add_rhs_matrix(linear);
      odeSolver.step_stage(0);
      odeSolver.final();
    }
    dg_solu.refine();
    dg_solu.copy_predict_to_ucoe();
    HyperbolicAlpt linear(dg_solu, oper_matx_alpt);
    lax_friedrichs(linear);
    const double h = 1. / std::pow(2., dg_solu.max_mesh_level());
    if (num_visc_elem != 0 && (viscosity_option == 1 || viscosity_option == 2)) {
      ArtificialViscosity artific_visc_operator(dg_solu, oper_matx_alpt, all_bas_alpt);
      artific_visc_operator.assemble_matrix(-artificial_viscosity_nu0 * h, viscosity_option);
      if (is_imex_option == 1) {
        IMEX43 odeSolver(artific_visc_operator, dt);
        odeSolver.init();
        for (int stage = 0; stage < odeSolver.num_stage; ++stage) {
          // the explicit part only for stages 2, 3, 4
          if (stage >= 2) {
            interpolate();
            dg_solu.set_rhs_zero();
            fast_rhs_herm.rhs_vol_scalar();
            fast_rhs_herm.rhs_flx_intp_scalar();
            odeSolver.set_rhs_zero();
            odeSolver.add_rhs_to_eigenvec();
            odeSolver.add_rhs_matrix(linear);
          }
          odeSolver.step_stage(stage);
          odeSolver.final();
          total_cg_iterations += odeSolver.last_cg_iterations;
        }
      } else {
        RK3SSP odeSolver(linear, dt);
        odeSolver.init();
        for (int stage = 0; stage < odeSolver.num_stage; ++stage) {
          interpolate();
          dg_solu.set_rhs_zero();
          fast_rhs_herm.rhs_vol_scalar();
          fast_rhs_herm.rhs_flx_intp_scalar();
          odeSolver.set_rhs_zero();
          odeSolver.add_rhs_to_eigenvec();
          odeSolver.add_rhs_matrix(linear);
          odeSolver.add_rhs_matrix(artific_visc_operator);
          odeSolver.step_stage(stage);
          odeSolver.final();
        }
      }
      dg_solu.coarsen();
      curr_time += dt;
      num_time_step++;
      continue;
    }
    RK3SSP odeSolver(linear, dt);
    odeSolver.init();
    for (int stage = 0; stage < odeSolver.num_stage; ++stage) {
      interpolate();
      dg_solu.set_rhs_zero();
      fast_rhs_herm.rhs_vol_scalar();
      fast_rhs_herm.rhs_flx_intp_scalar();
      if (num_visc_elem != 0 && viscosity_option == 3) {
        interp_lagr.support_gradu_Lagr_fast(dg_solu.viscosity_element, -artificial_viscosity_nu0 * h, is_intp_lagr,
                                            fast_lagr_intp, support_overlap == 1);
        fast_rhs_lagr.rhs_vol_scalar();
      }
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
            << " dof " << dg_solu.get_dof() << " checksum " << checksum << " viscosity elements " << total_visc_elem
            << " cg iterations " << total_cg_iterations << std::endl;
  if (!state_file.empty()) {
    IO inout(dg_solu);
    inout.output_state(state_file);
  }
  auto final_func = [&](std::vector<double> x) { return burgers.exact_2d(x[0], x[1], curr_time); };
  const std::vector<double> err = dg_solu.get_error_no_separable_scalar(final_func, num_gauss_pt_compute_error);
  std::cout << "L1, L2 and Linf error at final time: " << err[0] << ", " << err[1] << ", " << err[2] << std::endl;
  record_time.time("running time");
  return 0;
}
