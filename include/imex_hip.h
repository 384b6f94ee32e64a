/*
 * imex_hip.h -- C ABI of the MI355X (gfx950) IMEX-RK stage of IMEX_SfloW2D.
 *
 * Drop-in boundary for MODULE solver_2d of the reference (src/solver_2d.f90).
 * The Fortran driver keeps its time loop (src/IMEX_SfloW2D.f90:395-591) and
 * calls these entry points through ISO_C_BINDING.  All
 * state lives on the GPU in SoA layout; host arrays are borrowed only for the
 * duration of a call.  Arrays crossing the boundary use the reference's
 * Fortran order: q(n_vars, nx, ny) -> C index (k*nx + j)*n_vars + i.
 *
 * Return codes: 0 ok; > 0 physics error bit mask (the conditions on which the
 * reference prints and pauses, see IMEX_ERR_*); < 0 HIP / API error, text in
 * imex_last_error().  One host thread and one HIP stream per handle; calls
 * execute in order on that stream.
 */
#ifndef IMEX_HIP_H
#define IMEX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* physics error bits (reference pause / warning conditions) */
#define IMEX_ERR_NAN 1u          /* NaN after the RK stage, src/solver_2d.f90:1382-1397 */
#define IMEX_ERR_NEG_H 2u        /* q1 < -1e-7 (RK) or < -1e-10 (sedimentation), :1401-1438, :2514-2538 */
#define IMEX_ERR_NEG_SOLID 4u    /* solid mass <= -1e-7, :1440-1474, :2499-2511 */
#define IMEX_ERR_COLD 8u         /* T < 273 K after the RK stage, :1541-1560 */
#define IMEX_ERR_NEG_DEP 16u     /* negative deposition rate, src/constitutive_2d.f90:4860-4866 */
#define IMEX_ERR_SOLID_GT_TOT 32u /* sum of solid mass > q1 after sedimentation, :2546-2566 */
#define IMEX_WARN_RESCALE 64u    /* solid mass rescaled after the RK stage (reference warns, goes on) */

/* Run configuration: the namelist values of the configured path (gas
 * carrier, Voellmy or mu(I) friction, KT, minmod, IMEX RK2, Neumann boundaries).
 * Replaces the module variables read_param sets (src/inpout_2d.f90:986-5134). */
typedef struct imex_config {
  int32_t nx, ny;              /* comp_cells_x, comp_cells_y */
  int32_t n_solid;             /* 1..3 */
  int32_t energy_flag;
  int32_t settling_flag;
  int32_t entrainment_flag;
  int32_t radial_source_flag;
  int32_t n_thickness_levels;  /* 0..8 */
  int32_t n_dyn_pres_levels;   /* 0..8 */
  int32_t precision;           /* 0 = FP64 drop-in; 1 = FP32 form: state, right-hand side,
                                  reconstruction and fluxes in FP32, the Newton solve,
                                  thin-layer (h < 1e-3 m) primitive variables, deposit,
                                  maxima and the mass budget in FP64 (not in batches) */
  double dx;                   /* cell_size (dx = dy) */
  double x0, y0;
  double cfl, max_dt;
  double grav, mu, xi, pres, T_ambient, sp_heat_a, sp_gas_const_a, kin_visc_a;
  double rho_s[3], diam_s[3], sp_heat_s[3];
  double vel_source, T_source, h_source, time_param[4], alphas_source[3];
  double thickness_levels[8], dyn_pres_levels[8];
  /* erosion of an erodible layer (SOLID_TRANSPORT_PARAMETERS,
   * src/inpout_2d.f90:1813-1900); with erosion_coeff = 0 the others are
   * ignored as the reference ignores them.  erodible_fract as read_param
   * leaves them (summing to one; 1 for one class). */
  double erosion_coeff, erodible_porosity, T_erodible, erodible_fract[3];
  int32_t erodible_deposit_flag;
  int32_t topo_change_flag;    /* B_cent += dt * topo_term (src/solver_2d.f90:2493-2497) and the
                                  bed slopes recomputed after each device-driven step; needs
                                  imex_set_topography before the first step */
  /* friction law (RHEOLOGY_PARAMETERS): 0 or 1 = Voellmy-Salm (mu, xi);
   * 11 = mu(I) for dense granular flows, src/constitutive_2d.f90:4384-4458,
   * no implicit term; 12 = mu(I) plus the Voellmy turbulent term xi,
   * :4461-4537 and :3660-3683.  The mu(I) parameters must be >= 0
   * (src/inpout_2d.f90:3592-3630). */
  int32_t rheology_model;
  double mu_s, mu_2, muI_inf, I_0;
  /* stochastic friction (STOCHASTIC_PARAMETERS without transport of the
   * noise): each cell's xi is xi + Z(j,k) limited to [1, 1e5] in the
   * turbulent term of models 1 and 12 (src/constitutive_2d.f90:3663-3673);
   * Z is 0 until imex_set_noise gives it */
  int32_t stochastic_flag;
  /* pore pressure (PORE_PRESSURE_PARAMETERS, src/inpout_2d.f90:389-391,
   * :536-542, :681-710; maximum_solid_packing and sphericity_s from
   * SOLID_TRANSPORT_PARAMETERS, :819, :854): the conserved variable
   * q1 * p_exc after the solid classes (n_vars = 5 + n_solid), its implicit
   * decay (src/constitutive_2d.f90:3844-3993) with f_inhibit_mode 0 OFF,
   * 1 STEP, 2 STATIC, 3 DYNAMIC (smoothstep of order N_inh), the
   * Carman-Kozeny permeability when dynamic_permeability_flag, the gas loss
   * (:4887-5032) when gas_loss_flag, and the effective normal stress of the
   * friction laws.  Device-driven steps only (imex_run_steps); not with the
   * radial source, settling, entrainment, erosion, topography change, the
   * FP32 form, batches or the output kernels.  maximum_solid_packing 0
   * means the reference's default 0.64; sphericity_s 0 means 1. */
  int32_t pore_pressure_flag;
  int32_t gas_loss_flag;
  int32_t f_inhibit_mode;
  int32_t N_inh;
  int32_t dynamic_permeability_flag;
  double hydraulic_permeability, alpha_trans, maximum_solid_packing, sphericity_s[3];
  /* per-variable boundary conditions (WEST/EAST/SOUTH/NORTH_BOUNDARY_CONDITIONS,
   * src/solver_2d.f90:3190-3484), pore-pressure and transported-noise runs
   * only: with bc_custom,
   * bc_flag[8 side + var] and bc_value[8 side + var] for the sides W, E, S,
   * N and the conservative variables in order; flag 0 Dirichlet value, 1
   * slope value, 2 one-sided slope.  Without it every flag is 1 with value
   * 0 (the zero-gradient boundaries of the rest of the path). */
  int32_t bc_custom;
  int32_t bc_flag[32];
  double bc_value[32];
  /* the transported noise (STOCHASTIC_PARAMETERS stoch_transport_flag,
   * src/inpout_2d.f90:1466-1475; needs stochastic_flag): the conserved
   * variable q1 * Z after the solid classes and before the pore pressure
   * (n_vars = 4 + n_solid + 1 [+ 1]), advected with the flow
   * (src/constitutive_2d.f90:2428-2431, :2490-2493) and exchanged with the
   * mass (:5132-5137); the friction keeps reading the noise plane of
   * imex_set_noise.  The host gives its initial value q1 * Z
   * (getSteadyStateZ, src/stochastic_module.f90:110-126).  Device-driven
   * steps only, with the restrictions of pore-pressure runs and at most 8
   * variables. */
  int32_t stoch_transport_flag;
} imex_config;

/* Per-cell static fields computed once by the host (src/geometry_2d.f90:555-895
 * topography_reconstruction, :907-1323 init_source).  x index fastest,
 * nx*ny entries each.  source arrays may be NULL when there is no source. */
typedef struct imex_static {
  const double *B_prime_x;     /* B_prime_x_geom */
  const double *B_prime_y;     /* B_prime_y_geom */
  const int32_t *source_cell;  /* 0 / 1 (inside the source disc) / 2 (ring) */
  const int32_t *source_sides; /* bit0 sourceE, bit1 sourceW, bit2 sourceN, bit3 sourceS */
  const double *arc_perim;     /* cell_arc_perim */
  const double *arc_n_x;       /* cell_arc_n_x */
  const double *arc_n_y;       /* cell_arc_n_y */
} imex_static;

/* Display diagnostics of the driver's stdout line (src/IMEX_SfloW2D.f90:507-544),
 * full-grid compensated (Neumaier) sums. */
typedef struct imex_step_stats {
  double t, dt;
  double mass;        /* dx*dy*SUM(q(1,:,:)) */
  double volume;      /* dx*dy*SUM(qp(1,:,:)) over wet cells */
  double area;        /* dx*dy*COUNT(q(1,:,:) > 1e-7) */
  double solid_mass;  /* dx*dy*SUM(q(5:4+n_solid,:,:)) */
  int64_t active_cells; /* active cells summed over every check_solve since create */
  int64_t dried_cells; /* cells that went from wet to exactly dry since create */
  uint32_t err;        /* OR of physics error bits since the last call */
} imex_step_stats;

typedef struct imex_handle imex_handle;

/* download selection bits */
#define IMEX_DL_Q 1u
#define IMEX_DL_STATS 2u   /* hmax, pdynmax, mod_vel_max */
#define IMEX_DL_DEPOSIT 4u
#define IMEX_DL_VULN 8u
#define IMEX_DL_MASK 16u

/* Replaces allocate_solver_variables (src/solver_2d.f90:264-541): allocates
 * the device state for one run.  Device 0 unless IMEX_DEVICE is set. */
int imex_create(const imex_config *cfg, imex_handle **h);
/* Static fields (copied). */
int imex_set_static(imex_handle *h, const imex_static *s);
/* Initial or restart state, Fortran order q(n_vars,nx,ny); also sets t and dt
 * history and runs the initial active set and driver epilogue
 * (check_solve(.TRUE.), src/IMEX_SfloW2D.f90:265, :306-353). */
int imex_upload_state(imex_handle *h, const double *q_aos, double t, double dt0);
/* check_solve(solve_all), src/solver_2d.f90:667-858. */
int imex_check_solve(imex_handle *h, int solve_all);
/* timestep, src/solver_2d.f90:873-1006: the CFL dt (before the driver's
 * clipping).  Also evaluates the stage-1 explicit terms at time t.  The only
 * per-step blocking read-back (8 bytes). */
int imex_timestep(imex_handle *h, double t, double *dt_out);
/* imex_RK_solver, src/solver_2d.f90:1022-1634, with the driver's dt. */
int imex_rk_step(imex_handle *h, double dt);
/* update_erosion_deposition_cell(dt), src/solver_2d.f90:2382-2597. */
int imex_erosion_deposition(imex_handle *h, double dt);
/* driver epilogue (src/IMEX_SfloW2D.f90:449-505) and the stdout diagnostics. */
int imex_post_step(imex_handle *h, double t_new, imex_step_stats *out);
/* imex_post_step without any allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The diagnostics' per-block
 * partial sums (at most 512 blocks, the handle's fixed count) and their
 * combination live in a file-scope device array (32 KB per GPU, held until
 * the process ends) instead of the handle's buffer and host memory; the
 * result equals imex_post_step's bit for bit.  Calls that overlap on one GPU
 * share that array and race. */
int imex_post_step_noalloc(imex_handle *h, double t_new, imex_step_stats *out);
/* Up to nsteps full driver steps on the device, no host synchronisation
 * inside: check_solve, timestep, the driver's dt clip to t_end and growth
 * limit (src/IMEX_SfloW2D.f90:417-439), imex_RK_solver, sedimentation,
 * t += dt, epilogue.  The steps end once t reaches t_end (the driver loop's
 * condition t < t_end, :395): the remaining steps of the call do nothing,
 * so a call with t_end at the next output time stops exactly there.  dts
 * (may be NULL) receives each step's dt at the end (0 for a step not taken).
 * Returns the physics error bits raised by the call's own steps. */
int imex_run_steps(imex_handle *h, int nsteps, double t_end, double *dts);
/* imex_run_steps with HIP events around the kernel launches of every
 * sample_every-th step (steps n with n % sample_every == 0; an event record
 * costs the stream a few microseconds, so the other steps carry none) on the
 * handle's stream: kernel_ms[0..2] receive the summed durations over the
 * sampled steps of the active-set kernel, the flux kernel and the update
 * kernel (stage 2, sedimentation, epilogue), n_sampled their number;
 * cells[0] the active cells summed over all nsteps steps.  Has no
 * counterpart in the reference (measurement only). */
int imex_run_steps_timed(imex_handle *h, int nsteps, double t_end, double *dts,
                         int sample_every, double *kernel_ms, int *n_sampled, int64_t *cells);
/* Byte-counter calibration: copies n_doubles doubles between two device
 * buffers reps times (reads and writes 8*n_doubles bytes per launch). */
int imex_calib_copy(imex_handle *h, int64_t n_doubles, int reps);

/* current t and dt of the device-side driver */
int imex_get_time(imex_handle *h, double *t, double *dt);
/* set t and the dt history (dt_old, dt_old_old) of the device-side driver,
 * e.g. after host-driven steps or a restart */
int imex_set_time(imex_handle *h, double t, double dt_old, double dt_old_old);
/* t and the dt history of the device-side driver after device-driven steps
 * (the counterpart of imex_set_time) */
int imex_get_time_history(imex_handle *h, double *t, double *dt_old, double *dt_old_old);
/* active cells summed over every check_solve since create, and the
 * device-driven steps since the last upload (measurement; no counterpart in
 * the reference) */
int imex_get_counters(imex_handle *h, int64_t *active_cells, int64_t *steps);
/* cells the update kernel processed (its work list: active cells whose state
 * or stage-1 right-hand side is not all zero) summed over the steps since
 * create, and the cells of the always-active boundary frame that every
 * check_solve adds to the active count (src/solver_2d.f90:351-354) while no
 * kernel has work there (measurement; no counterpart in the reference) */
int imex_get_work_counters(imex_handle *h, int64_t *work_cells, int64_t *frame_cells);
/* bytes of device memory the handle holds now (its state, static fields,
 * lists, and any output staging allocated so far; no counterpart in the
 * reference: admission of served runs by footprint) */
int imex_get_footprint(imex_handle *h, size_t *bytes);
/* topography_reconstruction (src/geometry_2d.f90:555-895, called at
 * src/IMEX_SfloW2D.f90:445 when topo_change_flag): the bed slopes
 * B_prime_x_geom / B_prime_y_geom from the device's B_cent, for host-driven
 * steps (the device-driven steps do it themselves); no-op without
 * topo_change_flag. */
int imex_topography_reconstruction(imex_handle *h);
/* The erodible layer erodible(n_solid, nx, ny) (Fortran order; read_erodible
 * and src/inpout_2d.f90:1982-1986) of a run with erosion. */
int imex_set_erodible(imex_handle *h, const double *erodible);
/* The stochastic value Z(nx, ny) of every cell (Fortran order, x fastest) of
 * a run with stochastic_flag: the array solver_2d's Z holds after
 * getSteadyStateZ (src/stochastic_module.f90:79-132) or read_restart_file
 * (src/inpout_2d.f90:8153).  The value stays as given: without transport of
 * the noise the reference does not update Z during the run. */
int imex_set_noise(imex_handle *h, const double *Z);
/* erosion(nx, ny, n_solid) and erodible(n_solid, nx, ny) (either may be
 * NULL); zeros for a run without erosion. */
int imex_get_erosion(imex_handle *h, double *erosion, double *erodible);
/* Lazy host mirrors: copies the selected device arrays (any pointer may be
 * NULL when not selected).  q in Fortran order; deposit(nx,ny,n_solid);
 * vuln as one int32 per (level, cell) in Fortran order; mask one int32 per cell. */
int imex_download(imex_handle *h, unsigned what, double *q_aos, double *hmax,
                  double *pdynmax, double *modvelmax, double *deposit, int32_t *vuln,
                  int32_t *mask);
int imex_last_error(imex_handle *h, char *buf, size_t len);
void imex_destroy(imex_handle *h);

/* ---- device-side outputs ---------------------------------------------------
 * The grid work of the driver's periodic outputs, on the device state; the
 * caller formats and writes the files.  Primitive variables are those of the
 * driver's module array qp: qc_to_qp(q) where q(1) > 0, else zero with
 * T = T_ambient (src/IMEX_SfloW2D.f90:318-360, :456-503).  Cell indices are
 * 0-based, linear j + nx*k (Fortran array order). */

/* B_cent (nx*ny, x fastest): output_runout reads it when the runout origin
 * is not set (src/inpout_2d.f90:7293), the .p_2d records write it, and a run
 * with topo_change_flag evolves it (required there before the first step). */
int imex_set_topography(imex_handle *h, const double *B_cent);
int imex_get_topography(imex_handle *h, double *B_cent);

/* The whole-array expressions of one output_runout call
 * (src/inpout_2d.f90:7170-7427) over hpos = (qp(1) > 1e-5) .AND.
 * (SUM(qp(5:4+n_solid)) >= alphas_threshold): counts, sums, MAXVAL and
 * MAXLOC (ties to the first index, as MAXLOC).  dist, dist_x, dist_y are
 * 0 outside hpos.  hpos_old is a device mask that the call with
 * IMEX_RUNOUT_STORE overwrites with hpos after counting; it starts all
 * false (the reference leaves it uninitialised until its first call sets
 * it).  Sums are double-double accumulations rounded once (the reference
 * sums in array order). */
typedef struct imex_runout_red {
  double max_h;                     /* MAXVAL(qp(1,:,:)) */
  double sum_q1, sum_xq1, sum_yq1;  /* SUM(q(1,:,:)), SUM(X*q(1,:,:)), SUM(Y*q(1,:,:)) */
  int64_t n_hpos;                   /* COUNT(hpos) */
  int64_t n_hpos_old;               /* COUNT(hpos_old) */
  int64_t n_new;                    /* COUNT(hpos .AND. .NOT. hpos_old) */
  int64_t imax, imax_x, imax_y;     /* MAXLOC(dist), MAXLOC(dist_x), MAXLOC(dist_y) */
  double dist_max, dist_x_max, dist_y_max; /* the three maxima */
  double dist_at_imax_x, dist_at_imax_y;   /* dist(imax_x), dist(imax_y) (:7315, :7323) */
  int64_t ib;                       /* with IMEX_RUNOUT_B: MAXLOC of B_cent where
                                       qp(1) > 1e-5, else 0 (:7293-7294); -1 otherwise */
  double b_max;
} imex_runout_red;
#define IMEX_RUNOUT_STORE 1u  /* hpos_old := hpos after counting */
#define IMEX_RUNOUT_B 2u      /* also the MAXLOC of the wet cells' B_cent */
int imex_runout_reduce(imex_handle *h, double x0_runout, double y0_runout,
                       double alphas_threshold, unsigned flags, imex_runout_red *out);

/* output_probes (src/inpout_2d.f90:6926-7155) for n probes at xy[2*p],
 * xy[2*p+1]: per probe the values the reference writes after the time
 * column, in its order -- h, rho_m, T, u, v, alpha_s(1..n_solid) (h*alpha
 * over h), deposit(1..n_solid) when settling is on, p_dyn -- from the
 * bilinear interp_2d_scalarB (src/geometry_2d.f90:1535-1599); out holds
 * n * (6 + n_solid + settling*n_solid) values.  A point left of the first
 * cell centre reads 0 (the reference reads x_comp(0) there, outside the
 * array). */
int imex_probes(imex_handle *h, int n, const double *xy, double *out);

/* Snapshots for output_solution (src/inpout_2d.f90:5934-6423) and
 * write_restart_file (:8037-8084).  imex_snapshot_begin assembles the
 * selected records on the device (in stream order after the steps already
 * issued) and copies them to the caller's host buffers on a second stream;
 * it returns without waiting (except at the first snapshot of a kind, or of a
 * larger grid, which allocates its device staging), so the next steps run
 * while the copy (and the caller's writing of the previous snapshot)
 * proceeds.  The buffers must
 * stay valid, and be page-locked for the copy to overlap (imex_host_alloc),
 * until imex_snapshot_wait returns.  Like the reference, the .q_2d pass
 * zeroes |q| < 1e-20 in the device state and the .p_2d pass zeroes
 * |deposit| < 1e-20.  Records are one row per cell in Fortran order (x
 * fastest); imex_snapshot_cols gives the values per cell of one kind. */
#define IMEX_SNAP_Q 1u      /* q(n_vars, nx, ny), after the .q_2d zeroing when selected */
#define IMEX_SNAP_Q2D 2u    /* x, y, q(1:n_vars) (src/inpout_2d.f90:6043-6058) */
#define IMEX_SNAP_DEP2D 4u  /* x, y, deposit(j,k,1:n_solid) (:6071-6075) */
#define IMEX_SNAP_P2D 8u    /* x, y, h, u, v, B, h+B, alpha_s(n_solid), T, rho_m, g', deposit,
                               erosion, erodible, alpha_l, u*, Ri, Rouse(n_solid), mu_eff,
                               hmax, pdynmax, mod_vel_max, I (:6390-6396): 19 + 4 n_solid
                               values; B from imex_set_topography (0 without it) */
#define IMEX_SNAP_FIELDS 16u /* deposit(nx,ny,n_solid), hmax, pdynmax, mod_vel_max,
                               erosion(nx,ny,n_solid), erodible as [class][cell] planes,
                               B_cent (0 without imex_set_topography): 3 n_solid + 4 */
#define IMEX_SNAP_ERO2D 32u /* x, y, erosion(j,k,1:n_solid) (:6060-6063), erosion runs */
#define IMEX_SNAP_ERD2D 64u /* x, y, erodible(1:n_solid,j,k) (:6065-6066), erosion runs */
int imex_snapshot_cols(imex_handle *h, unsigned kind);
int imex_snapshot_begin(imex_handle *h, unsigned what, double *q, double *q2d, double *dep2d,
                        double *p2d, double *fields, double *ero2d, double *erd2d);
int imex_snapshot_wait(imex_handle *h);
/* page-locked host memory for snapshot buffers */
int imex_host_alloc(size_t bytes, void **p);
int imex_host_free(void *p);

/* Host writers (no device work).  imex_write_records writes n_rows records
 * of n_col values as the reference's formatted output: one line per record
 * with fmt IMEX_FMT_E20_12 ('(2e20.12,100(e20.12))', the .q_2d files) or
 * IMEX_FMT_ES15_7 ('(100ES15.7E2)', .p_2d), a list-directed blank record
 * after every nx records and two at the end; n_threads formatting threads
 * (0: one per host core).  imex_write_restart writes restart.bin as
 * write_restart_file's unformatted sequential records (4-byte markers;
 * records of 2 GiB or more are rejected; Z, when not NULL, is the
 * stochastic_flag record).  imex_format_value formats one
 * value (fmt also IMEX_FMT_ES14_7, the probe files' ES14.7E2) into out
 * (at least 21 bytes). */
#define IMEX_FMT_E20_12 0
#define IMEX_FMT_ES15_7 1
#define IMEX_FMT_ES14_7 2
int imex_write_records(const char *path, const double *rec, int64_t n_rows, int n_col, int nx,
                       int fmt, int n_threads);
int imex_write_restart(const char *path, int nx, int ny, int n_vars, int n_solid, double t,
                       double dt, const double *q_aos, const double *B_cent,
                       const double *erodible, const double *deposit, const double *erosion,
                       const double *hmax, const double *pdynmax, const double *mod_vel_max,
                       const double *Z);
int imex_format_value(int fmt, double x, char *out);

/* ---- batches of independent runs -----------------------------------------
 * No single counterpart in the reference: the reference runs an ensemble as
 * independent processes, one driver loop each (EXAMPLES/
 * EXAMPLE_ENSEMBLE_ETNA/launch_jobs.py).  A batch advances every member run
 * (a handle, created, given its static fields and uploaded as above) by
 * device-driven driver steps (imex_run_steps) in one launch per kernel for
 * all members; each member keeps its own dt, t and t_end and stops once t
 * reaches t_end.  A member's results equal those of imex_run_steps on the
 * same handle.  All members have the same n_solid and device. */
typedef struct imex_batch imex_batch;
int imex_batch_create(int n_solid, imex_batch **b);
/* as imex_batch_create, for members of the given imex_config.precision (0:
 * FP64, 1: the FP32 form); every member of a batch has its precision */
int imex_batch_create_form(int n_solid, int precision, imex_batch **b);
/* blocks per member and launch; 0 (the default): about 16 blocks per CU
 * shared by the members, 8 to 1024 per member */
int imex_batch_set_blocks(imex_batch *b, int blocks_per_member);
/* launch shape: flux_threads 256 (the default: each hot half tile of a
 * member on 256 threads, two passes per phase) or 512 (one pass per phase,
 * or two half tiles per block when a member has more than blocks); graphs 1
 * (the default): each step is a captured HIP graph (one of 8 steps and one
 * of 1 step, recaptured when the members or the launch shape change)
 * replayed by imex_batch_run, 0: the step's kernels are launched one by
 * one.  Results do not depend on either setting. */
int imex_batch_set_launch(imex_batch *b, int flux_threads, int graphs);
/* the handle joins at the next step and runs until its t reaches t_end; the
 * handle's own calls must not be used while it is a member */
int imex_batch_add(imex_batch *b, imex_handle *h, double t_end);
int imex_batch_remove(imex_batch *b, imex_handle *h);
/* nsteps steps of every member that has not reached its t_end (asynchronous) */
int imex_batch_run(imex_batch *b, int nsteps);
/* per member, in the order added: done flag (t reached t_end), device-driven
 * steps since its upload, and t; n = number of members; any output may be
 * NULL; synchronises */
int imex_batch_status(imex_batch *b, int n, int32_t *done, int32_t *steps, double *t);
int imex_batch_last_error(imex_batch *b, char *buf, size_t n);
void imex_batch_destroy(imex_batch *b);

/* Per-function batch entry points (device evaluation of single reference
 * routines on n host-provided states; for known-answer tests).  States are
 * n contiguous vectors of n_vars (qc) or n_vars+2 (qp) values. */
int imex_test_qc_to_qp(imex_handle *h, int n, const double *qc, double *qp, double *pdyn);
int imex_test_qp_to_qc(imex_handle *h, int n, const double *qp, double *qc);
int imex_test_fluxes(imex_handle *h, int n, int dir, const double *qc, const double *qp,
                     double *flux);
int imex_test_speeds(imex_handle *h, int n, int dir, const double *qp, double *out2);
int imex_test_implicit_cplx(imex_handle *h, int n, const double *qre, const double *qim,
                            double *ore, double *oim);
/* the implicit terms as the Newton residual evaluates them at real states
 * (src/solver_2d.f90:2271 calls src/constitutive_2d.f90:3513 with r_qj): a
 * real-arithmetic evaluation equal to the complex one's real part, with the
 * complex evaluation as fallback; fast[i] = 1 where the real one was used */
int imex_test_implicit_real(imex_handle *h, int n, const double *q, double *out, int *fast);
int imex_test_semi_impl(imex_handle *h, int n, const double *qp, double *si);
int imex_test_expl(imex_handle *h, int n, const double *bx, const double *by, double t,
                   const double *qp, double *expl);
int imex_test_settling(imex_handle *h, int n, const double *diam, const double *rhos,
                       const double *rhoc, const double *ikv, double *out);
int imex_test_mass_exchange_ers(imex_handle *h, int n, const double *qp, const double *erod,
                                double dt, double *dep, double *eq, double *ers);
int imex_test_mass_exchange(imex_handle *h, int n, const double *qp, double dt,
                            double *dep, double *eq);
int imex_test_newton(imex_handle *h, int n, double dt, const double *q_si, const double *q0,
                     const double *R, double *q, int32_t *iters);
/* correctly rounded arithmetic probes: out[4i..4i+3] = a/b, sqrt(a), csqrt(a+ib) */
int imex_test_arith(imex_handle *h, int n, const double *a, const double *b, double *out);
/* the path's x**y (correctly rounded, as the reference's libm pow is in all
 * but rare cases) on the device, and the same code compiled for the host CPU
 * (no GPU needed; for checking the algorithm on the CPU) */
int imex_test_pow(imex_handle *h, int n, const double *x, const double *y, double *out);
int imex_host_pow(int n, const double *x, const double *y, double *out);

#ifdef __cplusplus
}
#endif
#endif
