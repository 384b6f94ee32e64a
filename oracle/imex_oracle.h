/*
 * CPU restatement of the IMEX_SfloW2D IMEX-RK stage (test infrastructure).
 *
 * This is the oracle: tests/ and bench.py's cpu_baseline leg use it as the
 * checker for the HIP path.  It is never part of the product path.
 */
#ifndef IMEX_ORACLE_H
#define IMEX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Integer and real parameter vectors share the layout of oracle/refpy.py
 * pack_params(): ip[16], rp[73]. */
typedef struct oracle_s oracle_t;

oracle_t *oracle_create(const int *ip, const double *rp);
void oracle_destroy(oracle_t *o);

/* Static per-cell fields (x index fastest, nx*ny each).  sides bit 0..3 =
 * sourceE, sourceW, sourceN, sourceS. */
void oracle_set_static(oracle_t *o, const double *bprime_x, const double *bprime_y,
                       const int *source_cell, const int *sides,
                       const double *arc_perim, const double *arc_nx,
                       const double *arc_ny, double h_source);
/* q in Fortran order q(var, j, k); also runs check_solve(.TRUE.) and the
 * initial driver epilogue. */
void oracle_init_state(oracle_t *o, const double *q);

/* n full driver steps; dts[n], ncells[n] receive dt and the active count.
 * Returns the OR of the physics error bits raised. */
int oracle_run(oracle_t *o, int n, double *dts, int *ncells);

void oracle_get_q(const oracle_t *o, double *q);
void oracle_set_q(oracle_t *o, const double *q);
void oracle_get_fields(const oracle_t *o, double *hmax, double *pdynmax,
                       double *modvelmax, double *deposit, double *t, double *dt);
void oracle_get_vuln(const oracle_t *o, int *v);
void oracle_get_mask(const oracle_t *o, int *m);
/* x**y of the path: mode 0 (default) the C library's pow, as the reference
 * build calls it; mode 1 correctly rounded (quad-precision pow rounded once),
 * the rounding the HIP path computes */
void oracle_set_pow_mode(oracle_t *o, int mode);
double oracle_pow_cr(double x, double y);
/* driver time and dt history (restart from a saved state) */
void oracle_set_time(oracle_t *o, double t, double dt_old, double dt_old_old);
void oracle_set_t_end(oracle_t *o, double t_end);
void oracle_mass_exchange_ers(const oracle_t *o, int n, const double *qp, const double *erod,
                              double dt, double *dep, double *eq, double *ers);
void oracle_set_erodible(oracle_t *o, const double *e);
void oracle_set_noise(oracle_t *o, const double *Z);
void oracle_set_bcent(oracle_t *o, const double *b);
void oracle_get_bcent(const oracle_t *o, double *b);
void oracle_get_erosion(const oracle_t *o, double *ers, double *erd);
double oracle_get_t(const oracle_t *o);
/* number of cells that went from wet (q1 > 0) to exactly dry in a step, and
 * reads of a face or stage state left from an earlier step that differ from
 * the state a fresh evaluation gives (the reference's stale-state quirk) */
void oracle_get_counters(const oracle_t *o, long *dried, long *stale_reads);
void oracle_get_stale_cells(const oracle_t *o, unsigned char *out);
void oracle_set_fresh(oracle_t *o, int fresh);
/* first cell (j + nx*k) whose assembled state held a NaN in the last
 * oracle_run call (the reference's print-and-pause at
 * src/solver_2d.f90:1382-1397), -1 when none */
long oracle_get_nan_cell(const oracle_t *o);
/* watch the west and east x faces of cell (j, k) (0-based): at every step
 * that computes them, their KT fluxes before the x solid-flux cap
 * (src/solver_2d.f90:2903-2912) and after the zero-velocity rule are kept;
 * oracle_watched_faces copies them (pre[16], post[16]: face 0 then face 1,
 * n_vars values each in slots of 8) and returns bit 0 / bit 1 for the faces
 * computed since the watch was set */
void oracle_watch_cell(oracle_t *o, int j, int k);
int oracle_watched_faces(const oracle_t *o, double *pre, double *post);
/* work counters since oracle_work_reset (process-wide): out[18] = Newton
 * solves, Newton iterations, max iterations of one solve, residual
 * evaluations, Jacobians, line-search trials, max residual + Jacobian-column
 * evaluations of one cell's solve, settling-velocity calls, settling
 * iterations, max iterations of one settling call, and the histogram of
 * Newton iterations per solve (1, 2, 3, 4, 5-8, 9-16, 17-32, > 32) */
void oracle_work_reset(void);
void oracle_work_stats(long *out);

/* per-function entry points: n states of n_vars / n_vars+2 values each */
void oracle_qc_to_qp(const oracle_t *o, int n, const double *qc, double *qp, double *pdyn);
void oracle_qp_to_qc(const oracle_t *o, int n, const double *qp, double *qc);
void oracle_mixt_var(const oracle_t *o, int n, const double *qp, double *out6);
void oracle_fluxes(const oracle_t *o, int n, int dir, const double *qc, const double *qp,
                   double *flux);
void oracle_speeds(const oracle_t *o, int n, int dir, const double *qp, double *out2);
void oracle_implicit(const oracle_t *o, int n, const double *qc, double *nh);
void oracle_implicit_cplx(const oracle_t *o, int n, const double *qre, const double *qim,
                          double *ore, double *oim);
void oracle_semi_impl(const oracle_t *o, int n, const double *qc, const double *qp,
                      double *si);
void oracle_expl(const oracle_t *o, int n, const double *bx, const double *by, double t,
                 const double *qp, double *expl);
void oracle_settling(const oracle_t *o, int n, const double *diam, const double *rhos,
                     const double *rhoc, const double *ikv, double *out);
void oracle_mass_exchange(const oracle_t *o, int n, const double *qp, double dt,
                          double *dep, double *eq);
/* the limited slope of limit() with the minmod limiter from three values at
 * three abscissae (src/geometry_2d.f90:1776-1806, :1855-1876) */
double oracle_limit_minmod(const double *v3, const double *z3);
/* DGESV (n x n, one right-hand side) as the reference links it */
void oracle_dgesv(int n, double *a, double *b);
void oracle_get_stage(const oracle_t *o, int j, int k, double *out);
/* one Newton solve: q_si, q0, R (n_vars each) -> q (n_vars); returns iterations */
int oracle_newton(const oracle_t *o, double dt, const double *q_si, const double *q0,
                  const double *R, double *q);

#ifdef __cplusplus
}
#endif
#endif
