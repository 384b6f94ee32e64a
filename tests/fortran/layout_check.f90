! Prints C_SIZEOF and the byte offset of every component of the BIND(C)
! types of imex_hip_binding (compared with sizeof/offsetof of the C
! structs of include/imex_hip.h by tests/test_fortran_binding.py).
PROGRAM layout_check
  USE, INTRINSIC :: iso_c_binding
  USE imex_hip_binding
  IMPLICIT NONE
  TYPE(imex_config), TARGET :: c
  TYPE(imex_static), TARGET :: s
  TYPE(imex_step_stats), TARGET :: st
  TYPE(imex_runout_red), TARGET :: r
  INTEGER(c_intptr_t) :: base
  WRITE(*,'(A,1X,I0)') 'imex_config.size', C_SIZEOF(c)
  base = TRANSFER(C_LOC(c), base)
  WRITE(*,'(A,1X,I0)') 'imex_config.nx', TRANSFER(C_LOC(c%nx), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.ny', TRANSFER(C_LOC(c%ny), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.n_solid', TRANSFER(C_LOC(c%n_solid), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.energy_flag', TRANSFER(C_LOC(c%energy_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.settling_flag', TRANSFER(C_LOC(c%settling_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.entrainment_flag', TRANSFER(C_LOC(c%entrainment_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.radial_source_flag', TRANSFER(C_LOC(c%radial_source_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.n_thickness_levels', TRANSFER(C_LOC(c%n_thickness_levels), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.n_dyn_pres_levels', TRANSFER(C_LOC(c%n_dyn_pres_levels), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.precision', TRANSFER(C_LOC(c%precision), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.dx', TRANSFER(C_LOC(c%dx), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.x0', TRANSFER(C_LOC(c%x0), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.y0', TRANSFER(C_LOC(c%y0), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.cfl', TRANSFER(C_LOC(c%cfl), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.max_dt', TRANSFER(C_LOC(c%max_dt), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.grav', TRANSFER(C_LOC(c%grav), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.mu', TRANSFER(C_LOC(c%mu), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.xi', TRANSFER(C_LOC(c%xi), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.pres', TRANSFER(C_LOC(c%pres), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.T_ambient', TRANSFER(C_LOC(c%T_ambient), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.sp_heat_a', TRANSFER(C_LOC(c%sp_heat_a), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.sp_gas_const_a', TRANSFER(C_LOC(c%sp_gas_const_a), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.kin_visc_a', TRANSFER(C_LOC(c%kin_visc_a), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.rho_s', TRANSFER(C_LOC(c%rho_s), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.diam_s', TRANSFER(C_LOC(c%diam_s), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.sp_heat_s', TRANSFER(C_LOC(c%sp_heat_s), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.vel_source', TRANSFER(C_LOC(c%vel_source), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.T_source', TRANSFER(C_LOC(c%T_source), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.h_source', TRANSFER(C_LOC(c%h_source), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.time_param', TRANSFER(C_LOC(c%time_param), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.alphas_source', TRANSFER(C_LOC(c%alphas_source), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.thickness_levels', TRANSFER(C_LOC(c%thickness_levels), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.dyn_pres_levels', TRANSFER(C_LOC(c%dyn_pres_levels), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.erosion_coeff', TRANSFER(C_LOC(c%erosion_coeff), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.erodible_porosity', TRANSFER(C_LOC(c%erodible_porosity), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.T_erodible', TRANSFER(C_LOC(c%T_erodible), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.erodible_fract', TRANSFER(C_LOC(c%erodible_fract), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.erodible_deposit_flag', TRANSFER(C_LOC(c%erodible_deposit_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.topo_change_flag', TRANSFER(C_LOC(c%topo_change_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.rheology_model', TRANSFER(C_LOC(c%rheology_model), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.mu_s', TRANSFER(C_LOC(c%mu_s), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.mu_2', TRANSFER(C_LOC(c%mu_2), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.muI_inf', TRANSFER(C_LOC(c%muI_inf), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.I_0', TRANSFER(C_LOC(c%I_0), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.stochastic_flag', TRANSFER(C_LOC(c%stochastic_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.pore_pressure_flag', TRANSFER(C_LOC(c%pore_pressure_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.gas_loss_flag', TRANSFER(C_LOC(c%gas_loss_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.f_inhibit_mode', TRANSFER(C_LOC(c%f_inhibit_mode), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.N_inh', TRANSFER(C_LOC(c%N_inh), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.dynamic_permeability_flag', TRANSFER(C_LOC(c%dynamic_permeability_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.hydraulic_permeability', TRANSFER(C_LOC(c%hydraulic_permeability), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.alpha_trans', TRANSFER(C_LOC(c%alpha_trans), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.maximum_solid_packing', TRANSFER(C_LOC(c%maximum_solid_packing), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.sphericity_s', TRANSFER(C_LOC(c%sphericity_s), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.bc_custom', TRANSFER(C_LOC(c%bc_custom), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.bc_flag', TRANSFER(C_LOC(c%bc_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.bc_value', TRANSFER(C_LOC(c%bc_value), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_config.stoch_transport_flag', TRANSFER(C_LOC(c%stoch_transport_flag), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_static.size', C_SIZEOF(s)
  base = TRANSFER(C_LOC(s), base)
  WRITE(*,'(A,1X,I0)') 'imex_static.B_prime_x', TRANSFER(C_LOC(s%B_prime_x), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_static.B_prime_y', TRANSFER(C_LOC(s%B_prime_y), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_static.source_cell', TRANSFER(C_LOC(s%source_cell), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_static.source_sides', TRANSFER(C_LOC(s%source_sides), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_static.arc_perim', TRANSFER(C_LOC(s%arc_perim), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_static.arc_n_x', TRANSFER(C_LOC(s%arc_n_x), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_static.arc_n_y', TRANSFER(C_LOC(s%arc_n_y), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.size', C_SIZEOF(st)
  base = TRANSFER(C_LOC(st), base)
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.t', TRANSFER(C_LOC(st%t), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.dt', TRANSFER(C_LOC(st%dt), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.mass', TRANSFER(C_LOC(st%mass), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.volume', TRANSFER(C_LOC(st%volume), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.area', TRANSFER(C_LOC(st%area), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.solid_mass', TRANSFER(C_LOC(st%solid_mass), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.active_cells', TRANSFER(C_LOC(st%active_cells), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.dried_cells', TRANSFER(C_LOC(st%dried_cells), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_step_stats.err', TRANSFER(C_LOC(st%err), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.size', C_SIZEOF(r)
  base = TRANSFER(C_LOC(r), base)
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.max_h', TRANSFER(C_LOC(r%max_h), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.sum_q1', TRANSFER(C_LOC(r%sum_q1), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.sum_xq1', TRANSFER(C_LOC(r%sum_xq1), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.sum_yq1', TRANSFER(C_LOC(r%sum_yq1), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.n_hpos', TRANSFER(C_LOC(r%n_hpos), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.n_hpos_old', TRANSFER(C_LOC(r%n_hpos_old), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.n_new', TRANSFER(C_LOC(r%n_new), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.imax', TRANSFER(C_LOC(r%imax), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.imax_x', TRANSFER(C_LOC(r%imax_x), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.imax_y', TRANSFER(C_LOC(r%imax_y), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.dist_max', TRANSFER(C_LOC(r%dist_max), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.dist_x_max', TRANSFER(C_LOC(r%dist_x_max), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.dist_y_max', TRANSFER(C_LOC(r%dist_y_max), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.dist_at_imax_x', TRANSFER(C_LOC(r%dist_at_imax_x), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.dist_at_imax_y', TRANSFER(C_LOC(r%dist_at_imax_y), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.ib', TRANSFER(C_LOC(r%ib), base) - base
  WRITE(*,'(A,1X,I0)') 'imex_runout_red.b_max', TRANSFER(C_LOC(r%b_max), base) - base
END PROGRAM layout_check
