/* Prints sizeof and offsetof of every member of the C structs of
 * include/imex_hip.h (compared with the Fortran BIND(C) types by
 * tests/test_fortran_binding.py). */
#include <stddef.h>
#include <stdio.h>
#include "imex_hip.h"

int main(void) {
  printf("imex_config.size %zu\n", sizeof(imex_config));
  printf("imex_config.nx %zu\n", offsetof(imex_config, nx));
  printf("imex_config.ny %zu\n", offsetof(imex_config, ny));
  printf("imex_config.n_solid %zu\n", offsetof(imex_config, n_solid));
  printf("imex_config.energy_flag %zu\n", offsetof(imex_config, energy_flag));
  printf("imex_config.settling_flag %zu\n", offsetof(imex_config, settling_flag));
  printf("imex_config.entrainment_flag %zu\n", offsetof(imex_config, entrainment_flag));
  printf("imex_config.radial_source_flag %zu\n", offsetof(imex_config, radial_source_flag));
  printf("imex_config.n_thickness_levels %zu\n", offsetof(imex_config, n_thickness_levels));
  printf("imex_config.n_dyn_pres_levels %zu\n", offsetof(imex_config, n_dyn_pres_levels));
  printf("imex_config.precision %zu\n", offsetof(imex_config, precision));
  printf("imex_config.dx %zu\n", offsetof(imex_config, dx));
  printf("imex_config.x0 %zu\n", offsetof(imex_config, x0));
  printf("imex_config.y0 %zu\n", offsetof(imex_config, y0));
  printf("imex_config.cfl %zu\n", offsetof(imex_config, cfl));
  printf("imex_config.max_dt %zu\n", offsetof(imex_config, max_dt));
  printf("imex_config.grav %zu\n", offsetof(imex_config, grav));
  printf("imex_config.mu %zu\n", offsetof(imex_config, mu));
  printf("imex_config.xi %zu\n", offsetof(imex_config, xi));
  printf("imex_config.pres %zu\n", offsetof(imex_config, pres));
  printf("imex_config.T_ambient %zu\n", offsetof(imex_config, T_ambient));
  printf("imex_config.sp_heat_a %zu\n", offsetof(imex_config, sp_heat_a));
  printf("imex_config.sp_gas_const_a %zu\n", offsetof(imex_config, sp_gas_const_a));
  printf("imex_config.kin_visc_a %zu\n", offsetof(imex_config, kin_visc_a));
  printf("imex_config.rho_s %zu\n", offsetof(imex_config, rho_s));
  printf("imex_config.diam_s %zu\n", offsetof(imex_config, diam_s));
  printf("imex_config.sp_heat_s %zu\n", offsetof(imex_config, sp_heat_s));
  printf("imex_config.vel_source %zu\n", offsetof(imex_config, vel_source));
  printf("imex_config.T_source %zu\n", offsetof(imex_config, T_source));
  printf("imex_config.h_source %zu\n", offsetof(imex_config, h_source));
  printf("imex_config.time_param %zu\n", offsetof(imex_config, time_param));
  printf("imex_config.alphas_source %zu\n", offsetof(imex_config, alphas_source));
  printf("imex_config.thickness_levels %zu\n", offsetof(imex_config, thickness_levels));
  printf("imex_config.dyn_pres_levels %zu\n", offsetof(imex_config, dyn_pres_levels));
  printf("imex_config.erosion_coeff %zu\n", offsetof(imex_config, erosion_coeff));
  printf("imex_config.erodible_porosity %zu\n", offsetof(imex_config, erodible_porosity));
  printf("imex_config.T_erodible %zu\n", offsetof(imex_config, T_erodible));
  printf("imex_config.erodible_fract %zu\n", offsetof(imex_config, erodible_fract));
  printf("imex_config.erodible_deposit_flag %zu\n", offsetof(imex_config, erodible_deposit_flag));
  printf("imex_config.topo_change_flag %zu\n", offsetof(imex_config, topo_change_flag));
  printf("imex_config.rheology_model %zu\n", offsetof(imex_config, rheology_model));
  printf("imex_config.mu_s %zu\n", offsetof(imex_config, mu_s));
  printf("imex_config.mu_2 %zu\n", offsetof(imex_config, mu_2));
  printf("imex_config.muI_inf %zu\n", offsetof(imex_config, muI_inf));
  printf("imex_config.I_0 %zu\n", offsetof(imex_config, I_0));
  printf("imex_config.stochastic_flag %zu\n", offsetof(imex_config, stochastic_flag));
  printf("imex_static.size %zu\n", sizeof(imex_static));
  printf("imex_static.B_prime_x %zu\n", offsetof(imex_static, B_prime_x));
  printf("imex_static.B_prime_y %zu\n", offsetof(imex_static, B_prime_y));
  printf("imex_static.source_cell %zu\n", offsetof(imex_static, source_cell));
  printf("imex_static.source_sides %zu\n", offsetof(imex_static, source_sides));
  printf("imex_static.arc_perim %zu\n", offsetof(imex_static, arc_perim));
  printf("imex_static.arc_n_x %zu\n", offsetof(imex_static, arc_n_x));
  printf("imex_static.arc_n_y %zu\n", offsetof(imex_static, arc_n_y));
  printf("imex_step_stats.size %zu\n", sizeof(imex_step_stats));
  printf("imex_step_stats.t %zu\n", offsetof(imex_step_stats, t));
  printf("imex_step_stats.dt %zu\n", offsetof(imex_step_stats, dt));
  printf("imex_step_stats.mass %zu\n", offsetof(imex_step_stats, mass));
  printf("imex_step_stats.volume %zu\n", offsetof(imex_step_stats, volume));
  printf("imex_step_stats.area %zu\n", offsetof(imex_step_stats, area));
  printf("imex_step_stats.solid_mass %zu\n", offsetof(imex_step_stats, solid_mass));
  printf("imex_step_stats.active_cells %zu\n", offsetof(imex_step_stats, active_cells));
  printf("imex_step_stats.dried_cells %zu\n", offsetof(imex_step_stats, dried_cells));
  printf("imex_step_stats.err %zu\n", offsetof(imex_step_stats, err));
  printf("imex_runout_red.size %zu\n", sizeof(imex_runout_red));
  printf("imex_runout_red.max_h %zu\n", offsetof(imex_runout_red, max_h));
  printf("imex_runout_red.sum_q1 %zu\n", offsetof(imex_runout_red, sum_q1));
  printf("imex_runout_red.sum_xq1 %zu\n", offsetof(imex_runout_red, sum_xq1));
  printf("imex_runout_red.sum_yq1 %zu\n", offsetof(imex_runout_red, sum_yq1));
  printf("imex_runout_red.n_hpos %zu\n", offsetof(imex_runout_red, n_hpos));
  printf("imex_runout_red.n_hpos_old %zu\n", offsetof(imex_runout_red, n_hpos_old));
  printf("imex_runout_red.n_new %zu\n", offsetof(imex_runout_red, n_new));
  printf("imex_runout_red.imax %zu\n", offsetof(imex_runout_red, imax));
  printf("imex_runout_red.imax_x %zu\n", offsetof(imex_runout_red, imax_x));
  printf("imex_runout_red.imax_y %zu\n", offsetof(imex_runout_red, imax_y));
  printf("imex_runout_red.dist_max %zu\n", offsetof(imex_runout_red, dist_max));
  printf("imex_runout_red.dist_x_max %zu\n", offsetof(imex_runout_red, dist_x_max));
  printf("imex_runout_red.dist_y_max %zu\n", offsetof(imex_runout_red, dist_y_max));
  printf("imex_runout_red.dist_at_imax_x %zu\n", offsetof(imex_runout_red, dist_at_imax_x));
  printf("imex_runout_red.dist_at_imax_y %zu\n", offsetof(imex_runout_red, dist_at_imax_y));
  printf("imex_runout_red.ib %zu\n", offsetof(imex_runout_red, ib));
  printf("imex_runout_red.b_max %zu\n", offsetof(imex_runout_red, b_max));
  return 0;
}
