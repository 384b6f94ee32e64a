/*
 * CPU restatement of the IMEX_SfloW2D IMEX-RK stage -- the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this (as the checker); the product path
 * never calls it.
 *
 * It restates, for the configured branches (gas carrier, no extra gases, no
 * liquid, alpha_flag=F, Voellmy rheology_model=1 or the mu(I) models 11
 * and 12 without pore pressure, KT fluxes, minmod,
 * IMEX RK2, Neumann boundaries, optional lateral radial source, settling and
 * air entrainment), the reference routines named at each function below, in
 * the reference's own data layout (Fortran order, variable index fastest),
 * with the same floating-point operation order:
 *   - left-to-right evaluation of * and / chains, x**2 as x*x, x**4 as
 *     x*(x*(x*x)), SUM / DOT_PRODUCT / MATMUL as sequential sums from 0;
 *   - complex arithmetic as the reference build evaluates it: products
 *     (ac-bd, ad+bc), real operands promoted to (r,0), DOT_PRODUCT of a complex
 *     first argument conjugated, division by the C99 Annex G scaled algorithm,
 *     square root by the C library csqrt;
 *   - the persistent module arrays, so a cell outside the active set keeps the
 *     face and stage states it had when it was last active (the reference's
 *     stale-state behaviour).
 *
 * Parity: pinned against oracle/_ref (the reference compiled from its own
 * sources) by tests/test_oracle_vs_ref.py, bit-for-bit per function and over
 * multi-step runs, and against closed forms in tests/test_oracle_closed_form.py.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <quadmath.h>

#include "imex_oracle.h"

#define EPS 2.220446049250313e-16 /* EPSILON(1.0_wp) */

typedef struct {
    double re, im;
} cx;

struct oracle_s {
    int nx, ny, ns, nv;
    int energy, settling, entrainment, radial, collapse, limiter, bcflag;
    int n_thk, n_pdyn;
    double dx, dy, dx2, dy2, one_by_dx, one_by_dy, x0, y0;
    double cfl, dt0, max_dt, t_start, t_end;
    double grav, mu, xi, pres, inv_pres, T_amb, cp_a, R_a, kin_visc_c, rho_a_amb;
    double eps_sing, eps_sing4;
    double rho_s[3], inv_rho_s[3], diam_s[3], cp_s[3], alphas_collapse[3], alphas_source[3];
    double rho_c_sub, T_erodible;
    /* erosion (src/inpout_2d.f90:1813-1900) */
    double erosion_coeff, erodible_porosity, coeff_porosity, erodible_fract[3];
    int erodible_deposit;
    int topo_change;        /* topo_change_flag */
    int rheology;           /* rheology_model: 1, 11 or 12 */
    double mu_s, mu_2, muI_inf, I_0;  /* mu(I) parameters (RHEOLOGY_PARAMETERS) */
    int stochastic;         /* stochastic_flag (no transport of the noise) */
    double *Z;              /* solver_2d's Z(nx, ny): the cells' stochastic values */
    double xi_cell;         /* xi_temp of the cell being solved (xi outside a step) */
    double *bcent;          /* B_cent (topography change) */
    double vel_source, T_source, h_source, time_param[4], pi_g;
    double thk[8], pdyn[8];
    double hcs, one_by_hcs; /* complex step h = n_vars*eps (src/solver_2d.f90:274-275) */
    double *x_comp, *x_stag, *y_comp, *y_stag;
    double *bx, *by, *arcp, *arcx, *arcy;
    int *scell;
    unsigned char *sE, *sW, *sN, *sS;
    /* module arrays of solver_2d */
    double *q, *qp, *q0, *q_fv;
    double *q_rk, *qp_rk; /* [stage] */
    double *divF, *NH, *SI, *expl; /* [stage] */
    double *qL, *qR, *qB, *qT, *qpL, *qpR, *qpB, *qpT;
    double *axn, *axp, *ayn, *ayp, *axmax, *aymax;
    double *Hx, *Hy;
    unsigned char *mask, *mtmp, *mx, *my;
    int *jc, *kc, *jsx, *ksx, *jsy, *ksy;
    int ncells, nfx, nfy;
    double *hmax, *pdynmax, *modvelmax, *deposit;
    double *erodible, *erosion; /* [class][cell] */
    unsigned char *vuln;
    double t, dt, dt_old, dt_old_old;
    long dried, stale;
    int fresh;              /* 1: states left by an earlier step read as a fresh evaluation */
    int err;
    long nan_cell;          /* first cell (j + nx k) whose assembled state held a NaN, -1 none */
    /* the two x faces of a watched cell: KT fluxes before the solid-flux cap
     * and after the zero-velocity rule, at the last step that computed them */
    int watch_j, watch_k, watch_seen[2];
    double watch_pre[2][8], watch_post[2][8];
    /* reconstruction-call stamps of the face sides and of the cell primitive
     * states, used only to count reads of states left from an earlier call */
    int call_id;
    int *stL, *stR, *stB, *stT, *stQP;
    unsigned char *stale_cell;  /* cells next to a stale read that differs (cumulative) */
    /* 0: x**y by the C library's pow, as the reference build calls it;
     * 1: x**y correctly rounded (quad-precision powq rounded once) */
    int pow_mode;
    /* pore pressure (PORE_PRESSURE_PARAMETERS, src/inpout_2d.f90:389-391,
     * :536-542, :681-710; maximum_solid_packing and sphericity_s from
     * SOLID_TRANSPORT_PARAMETERS): the conserved variable q1 * p_exc at
     * index ipp = 4 + ns + stoch_tr (0-based) */
    int pore, gas_loss, finh, N_inh, dyn_perm, ipp;  /* finh: 0 OFF 1 STEP 2 STATIC 3 DYNAMIC */
    /* the transported noise (stoch_transport_flag, src/inpout_2d.f90:642-644,
     * :1466-1475): the conserved variable q1 * Z at index ist = 4 + ns */
    int stoch_tr, ist;
    double hyd_perm, alpha_trans, max_pack, sph_s[3], pascal[64];
    /* per-variable boundary conditions (bcW/bcE/bcS/bcN of the
     * *_BOUNDARY_CONDITIONS namelists, src/solver_2d.f90:3190-3484):
     * side 0 W, 1 E, 2 S, 3 N; flag 0 Dirichlet value, 1 slope value,
     * 2 one-sided slope.  Without bc_custom every flag is 1 with value 0. */
    int bc_custom, bcf[4][8];
    double bcv[4][8];
};

/* Work counters of the stage-2 Newton solves and the settling iterations
 * since the last oracle_work_reset (process-wide; the oracle runs one thread):
 * they describe how long each cell's dependency chain is, which sets the
 * duration of a one-thread-per-cell update kernel. */
static struct {
    long newton_calls, newton_its, newton_max_its;
    long evalf, jac, ls_trials;
    long cell_evals, cell_evals_max;   /* residual + 2 complex Jacobian columns per cell */
    long settle_calls, settle_its, settle_max_its;
    long hist[8];                      /* newton iterations: 1,2,3,4,5-8,9-16,17-32,>32 */
} g_wk;

/* x**y correctly rounded: powq in binary128 (113-bit significand) rounded
 * once to double; exact unless x**y lies within 2**-58 ulp of a rounding
 * boundary of the double result.  powl settles the rounding first where it
 * can (see below): the same result, faster. */
double oracle_pow_cr(double x, double y)
{
    /* fast path: powl (64-bit significand, error within one long-double
     * ulp, i.e. 2**-11 ulp of the double result) decides the rounding unless
     * it lies within 2**-8 ulp of a rounding boundary; powq otherwise */
    const long double r = powl((long double)x, (long double)y);
    const double d = (double)r;
    if (isnormal(d) && isfinite(d)) {
        const long double up = (long double)nextafter(d, INFINITY) - (long double)d;
        const long double dn = (long double)d - (long double)nextafter(d, -INFINITY);
        const long double m = (r >= (long double)d) ? (long double)d + 0.5L * up
                                                    : (long double)d - 0.5L * dn;
        const long double ulp = up < dn ? up : dn;
        if (fabsl(r - m) > ldexpl(ulp, -8)) return d;
    }
    return (double)powq((__float128)x, (__float128)y);
}

/* REAL**REAL of the path: src/constitutive_2d.f90:4851, :5049, :5378 */
static inline double rpow(const oracle_t *o, double x, double y)
{
    return o->pow_mode ? oracle_pow_cr(x, y) : pow(x, y);
}

/* The other C-library functions the pore-pressure terms reach (through
 * the C library's cpow, which Fortran COMPLEX**REAL calls), each in the
 * mode of x**y: the C library's own (= the reference) or correctly rounded
 * (binary128 evaluation rounded once, the rounding the HIP path computes). */
static inline double r_log(const oracle_t *o, double x) { return o->pow_mode ? (double)logq(x) : log(x); }
static inline double r_log1p(const oracle_t *o, double x) { return o->pow_mode ? (double)log1pq(x) : log1p(x); }
static inline double r_exp(const oracle_t *o, double x) { return o->pow_mode ? (double)expq(x) : exp(x); }
static inline double r_atan2(const oracle_t *o, double y, double x)
{
    return o->pow_mode ? (double)atan2q(y, x) : atan2(y, x);
}
static inline double r_hypot(const oracle_t *o, double x, double y)
{
    return o->pow_mode ? (double)hypotq(x, y) : hypot(x, y);
}

/* ---------------------------------------------------------------------- */
/* indexing: cells c = j + nx*k (0-based), x faces j + (nx+1)*k,          */
/* y faces j + nx*k with k in 0..ny                                         */
/* ---------------------------------------------------------------------- */
#define CELL(o, j, k) ((j) + (o)->nx * (k))
#define FX(o, j, k) ((j) + ((o)->nx + 1) * (k))
#define FY(o, j, k) ((j) + (o)->nx * (k))

/* Fortran MIN(a, b) / MAX(a, b) as the reference build (amdflang, IEEE
 * flags) evaluates them: (a < b) ? a : b and (a > b) ? a : b, so a NaN or a
 * signed-zero tie returns the second argument (checked on MIN(1e15, NaN),
 * MIN(NaN, 1e15), MIN(0, -0), MIN(-0, 0) and the MAX counterparts). */
static inline double fmin_f(double a, double b) { return (a < b) ? a : b; }
static inline double fmax_f(double a, double b) { return (a > b) ? a : b; }

/* ---------------------------------------------------------------------- */
/* complex arithmetic as evaluated by the reference build                  */
/* ---------------------------------------------------------------------- */
static inline cx C(double re, double im) { cx z = {re, im}; return z; }
static inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cconj(cx a) { return C(a.re, -a.im); }
/* z**2 (src: integer power 2 of a complex) */
static inline cx csq(cx a) { return C(a.re * a.re - a.im * a.im, a.re * a.im + a.re * a.im); }
/* C99 Annex G.5.1 scaled complex division */
static cx cdiv(cx z, cx w)
{
    double a = z.re, b = z.im, c = w.re, d = w.im;
    int ilogbw = 0;
    double logbw = logb(fmax(fabs(c), fabs(d)));
    if (isfinite(logbw)) {
        ilogbw = (int)logbw;
        c = scalbn(c, -ilogbw);
        d = scalbn(d, -ilogbw);
    }
    double denom = c * c + d * d;
    double x = scalbn((a * c + b * d) / denom, -ilogbw);
    double y = scalbn((b * c - a * d) / denom, -ilogbw);
    if (isnan(x) && isnan(y)) {
        if ((denom == 0.0) && (!isnan(a) || !isnan(b))) {
            x = copysign(INFINITY, c) * a;
            y = copysign(INFINITY, c) * b;
        } else if ((isinf(a) || isinf(b)) && isfinite(c) && isfinite(d)) {
            a = copysign(isinf(a) ? 1.0 : 0.0, a);
            b = copysign(isinf(b) ? 1.0 : 0.0, b);
            x = INFINITY * (a * c + b * d);
            y = INFINITY * (b * c - a * d);
        } else if (isinf(logbw) && logbw > 0.0 && isfinite(a) && isfinite(b)) {
            c = copysign(isinf(c) ? 1.0 : 0.0, c);
            d = copysign(isinf(d) ? 1.0 : 0.0, d);
            x = 0.0 * (a * c + b * d);
            y = 0.0 * (b * c - a * d);
        }
    }
    return C(x, y);
}
static inline cx csqrt_c(cx a)
{
    double complex z = csqrt(CMPLX(a.re, a.im));
    return C(creal(z), cimag(z));
}

/* Fortran COMPLEX**REAL: the reference build calls the C library's
 * cpow(z, c + 0i) (checked on amdflang's output for z**3.0_wp and z**r).
 * The C library computes it as cexp(c * clog(z)); in the correctly rounded
 * mode the same steps run with correctly rounded log1p / log / hypot /
 * atan2 / exp / sin / cos: clog's real part by its case split (|x| = 1,
 * 1 < |x| < 2, 0.5 <= |x| < 1 with a tiny or a larger |y|, otherwise
 * log(hypot)), the product with c + 0i as (cL - 0 th, c th + 0 L), cexp as
 * exp(a) (cos b, sin b), with (cos b, sin b) = (1, b) for |b| <= DBL_MIN.
 * Finite arguments with |z| well inside the double range only (the
 * solver's porosities and thicknesses). */
static cx cpow_r(const oracle_t *o, cx z, double c)
{
    if (!o->pow_mode) {
        double complex r = cpow(CMPLX(z.re, z.im), CMPLX(c, 0.0));
        return C(creal(r), cimag(r));
    }
    double ax = fabs(z.re), ay = fabs(z.im);
    if (ax < ay) { double t = ax; ax = ay; ay = t; }
    double L;
    if (ax == 1.0) {
        L = r_log1p(o, ay * ay) / 2.0;
    } else if (ax > 1.0 && ax < 2.0 && ay < 1.0) {
        double d2m1 = (ax - 1.0) * (ax + 1.0);
        if (ay >= EPS) d2m1 += ay * ay;
        L = r_log1p(o, d2m1) / 2.0;
    } else if (ax < 1.0 && ax >= 0.5 && ay < EPS / 2.0) {
        double d2m1 = (ax - 1.0) * (ax + 1.0);
        L = r_log1p(o, d2m1) / 2.0;
    } else if (ax < 1.0 && ax >= 0.5 && ax * ax + ay * ay >= 0.5) {
        __float128 X = ax, Y = ay;
        L = r_log1p(o, (double)(X * X + Y * Y - 1)) / 2.0;
    } else {
        L = r_log(o, r_hypot(o, ax, ay)) - 0.0 * 0x1.62e42fefa39efp-1;
    }
    double th = r_atan2(o, z.im, z.re);
    double a = c * L - 0.0 * th, b = c * th + 0.0 * L;
    double sb, cb;
    if (fabs(b) > 2.2250738585072014e-308) {
        sb = (double)sinq(b);
        cb = (double)cosq(b);
    } else {
        sb = b;
        cb = 1.0;
    }
    double E = r_exp(o, a);
    return C(E * cb, E * sb);
}

/* ====================================================================== */
/* constitutive_2d                                                         */
/* ====================================================================== */

/* r_phys_var, configured branch: src/constitutive_2d.f90:550-976
 * (:642-697 mass fractions, :714-731 mixture heat, :754-771 temperature,
 *  :775-797 carrier density, :799-830 mixture, :944-972 velocity, p_dyn) */
static void r_phys_var(const oracle_t *o, const double *qc, double *h, double *u, double *v,
                       double *alphas, double *rho_m, double *T, double *red_grav,
                       double *p_dyn, double *pexc, double *zs)
{
    const int ns = o->ns;
    double xs[3], inv_qj1;
    if (qc[0] > EPS) {
        inv_qj1 = 1.0 / qc[0];
        for (int i = 0; i < ns; i++) xs[i] = qc[4 + i] * inv_qj1;
        double s = 0.0;
        for (int i = 0; i < ns; i++) s += qc[4 + i];
        if (s == qc[0]) {
            double sx = 0.0;
            for (int i = 0; i < ns; i++) sx += xs[i];
            for (int i = 0; i < ns; i++) xs[i] = xs[i] / sx;
        }
        /* transported noise and excess pore pressure:
         * src/constitutive_2d.f90:657, :659 */
        *zs = o->stoch_tr ? qc[o->ist] * inv_qj1 : 0.0;
        *pexc = o->pore ? qc[o->ipp] * inv_qj1 : 0.0;
    } else {
        *h = 0.0; *u = 0.0; *v = 0.0; *pexc = 0.0; *zs = 0.0;
        for (int i = 0; i < ns; i++) alphas[i] = 0.0;
        *rho_m = o->rho_a_amb;
        *T = o->T_amb;
        *red_grav = 0.0;
        *p_dyn = 0.0;
        return;
    }
    double xs_tot = 0.0;
    for (int i = 0; i < ns; i++) xs_tot += xs[i];
    double xc = 1.0 - xs_tot;
    double cpc_by_xc = (xc - 0.0) * o->cp_a + 0.0;
    double dot = 0.0;
    for (int i = 0; i < ns; i++) dot += xs[i] * o->cp_s[i];
    double cp_mix = dot + cpc_by_xc;
    double TT;
    if (qc[0] > o->eps_sing) {
        if (o->energy)
            TT = (qc[3] - 0.5 * (qc[1] * qc[1] + qc[2] * qc[2]) * inv_qj1) / (qc[0] * cp_mix);
        else
            TT = qc[3] / (qc[0] * cp_mix);
        if (TT <= 0.0) TT = o->T_amb;
    } else {
        TT = o->T_amb;
    }
    double Rc;
    if (xc > EPS)
        Rc = ((xc - 0.0) * o->R_a + 0.0) / xc;
    else
        Rc = o->R_a;
    double inv_rho_c = Rc * TT * o->inv_pres;
    double inv_rhom = 0.0;
    for (int i = 0; i < ns; i++) inv_rhom += xs[i] * o->inv_rho_s[i];
    inv_rhom = inv_rhom + xc * inv_rho_c;
    double rm = 1.0 / inv_rhom;
    for (int i = 0; i < ns; i++) alphas[i] = xs[i] * rm * o->inv_rho_s[i];
    double hh = qc[0] * inv_rhom;
    *red_grav = (rm - o->rho_a_amb) * inv_rhom * o->grav;
    double uu, vv;
    if (qc[0] > o->eps_sing) {
        uu = qc[1] * inv_qj1;
        vv = qc[2] * inv_qj1;
    } else {
        double q1 = qc[0];
        double den = sqrt(q1 * (q1 * (q1 * q1)) + o->eps_sing4);
        uu = 1.4142135623730951 * q1 * qc[1] / den;
        vv = 1.4142135623730951 * q1 * qc[2] / den;
    }
    *p_dyn = 0.5 * rm * (uu * uu + vv * vv);
    *h = hh; *u = uu; *v = vv; *rho_m = rm; *T = TT;
}

/* qc_to_qp: src/constitutive_2d.f90:1637-1690 */
static void qc_to_qp(const oracle_t *o, const double *qc, double *qp, double *p_dyn)
{
    double h, u, v, alphas[3], rho_m, T, rg, pexc, zs;
    r_phys_var(o, qc, &h, &u, &v, alphas, &rho_m, &T, &rg, p_dyn, &pexc, &zs);
    const int nv = o->nv;
    qp[0] = h;
    qp[1] = h * u;
    qp[2] = h * v;
    qp[3] = T;
    for (int i = 0; i < o->ns; i++) qp[4 + i] = alphas[i] * h;
    if (o->stoch_tr) qp[o->ist] = zs;  /* :1681 */
    if (o->pore) qp[o->ipp] = pexc;  /* :1683 */
    qp[nv] = u;
    qp[nv + 1] = v;
}

/* qp_to_qc, configured branch: src/constitutive_2d.f90:1718-2149
 * (:1791-1800 dry, :1809-1861 densities, :1923-1984 mixture, :2098-2146 qc) */
static void qp_to_qc(const oracle_t *o, const double *qp, double *qc)
{
    const int ns = o->ns, nv = o->nv;
    double h = qp[0];
    if (!(h > EPS)) {
        for (int i = 0; i < nv; i++) qc[i] = 0.0;
        return;
    }
    double hu = qp[1], hv = qp[2], u = qp[nv], v = qp[nv + 1], T = qp[3];
    double alphas[3], ar[3];
    for (int i = 0; i < ns; i++) alphas[i] = qp[4 + i] / qp[0];
    double at = 0.0;
    for (int i = 0; i < ns; i++) at += alphas[i];
    for (int i = 0; i < ns; i++) ar[i] = alphas[i] * o->rho_s[i];
    double rho_a = o->pres / (o->R_a * T);
    double rho_c = (((1.0 - 0.0) - at - 0.0) * rho_a + 0.0) / ((1.0 - 0.0) - at);
    double alphac = 1.0 - at;
    double sar = 0.0;
    for (int i = 0; i < ns; i++) sar += ar[i];
    double rho_m = alphac * rho_c + sar;
    double inv_rhom = 1.0 / rho_m;
    double xs[3];
    for (int i = 0; i < ns; i++) xs[i] = ar[i] * inv_rhom;
    double xc = alphac * rho_c * inv_rhom;
    double cp_c;
    if (xc > EPS)
        cp_c = ((xc - 0.0) * o->cp_a + 0.0) / xc;
    else
        cp_c = o->cp_a;
    double dot = 0.0;
    for (int i = 0; i < ns; i++) dot += xs[i] * o->cp_s[i];
    double cp_mix = dot + xc * cp_c;
    qc[0] = rho_m * h;
    qc[1] = rho_m * hu;
    qc[2] = rho_m * hv;
    if (o->energy) {
        if (h > 0.0)
            qc[3] = h * rho_m * (cp_mix * T + 0.5 * (u * u + v * v));
        else
            qc[3] = 0.0;
    } else {
        qc[3] = h * rho_m * cp_mix * T;
    }
    for (int i = 0; i < ns; i++) qc[4 + i] = xs[i] * qc[0];
    if (o->stoch_tr) qc[o->ist] = qp[o->ist] * qc[0];  /* :1979, :2141 */
    if (o->pore) qc[o->ipp] = qp[o->ipp] * qc[0];  /* :1981, :2143 */
}

/* eval_sp_heat, gas without liquid: src/constitutive_2d.f90:1513-1607 (:1568-1604) */
static void eval_sp_heat(const oracle_t *o, const double *alphas, double inv_rhom,
                         double *cp_c, double *cp_mix)
{
    const int ns = o->ns;
    double xs[3];
    for (int i = 0; i < ns; i++) xs[i] = alphas[i] * o->rho_s[i] * inv_rhom;
    double sxs = 0.0;
    for (int i = 0; i < ns; i++) sxs += xs[i];
    double xc = 1.0 - sxs;
    double c = 0.0;
    if (xc > EPS)
        c = ((xc - 0.0) * o->cp_a + 0.0) / xc;
    else
        c = o->cp_a;
    double dot = 0.0;
    for (int i = 0; i < ns; i++) dot += xs[i] * o->cp_s[i];
    *cp_c = c;
    *cp_mix = dot + xc * c;
}

/* mixt_var: src/constitutive_2d.f90:1342-1492 */
static void mixt_var(const oracle_t *o, const double *qp, double *Ri, double *rho_m,
                     double *rho_c, double *red_grav, int sp_heat_flag, double *cp_c,
                     double *cp_mix)
{
    const int ns = o->ns, nv = o->nv;
    double h = qp[0];
    if (qp[0] <= EPS) {
        *red_grav = 0.0;
        *rho_m = o->rho_a_amb;
        *rho_c = o->rho_a_amb;
        *cp_c = o->cp_a;
        *cp_mix = o->cp_a;
        *Ri = 0.0;
        return;
    }
    double u = qp[nv], v = qp[nv + 1], T = qp[3];
    double alphas[3];
    for (int i = 0; i < ns; i++) alphas[i] = qp[4 + i] / qp[0];
    double at = 0.0;
    for (int i = 0; i < ns; i++) at += alphas[i];
    double rho_a = o->pres / (o->R_a * T);
    double rc = (((1.0 - 0.0) - at - 0.0) * rho_a + 0.0) / ((1.0 - 0.0) - at);
    double dot = 0.0;
    for (int i = 0; i < ns; i++) dot += alphas[i] * o->rho_s[i];
    double rm = (1.0 - at) * rc + dot;
    double inv_rhom = 1.0 / rm;
    double rg = (rm - o->rho_a_amb) / rm * o->grav;
    double u2 = u * u + v * v;
    if (u2 > 0.0)
        *Ri = fmin_f(1.0e15, rg * h / u2);
    else
        *Ri = 0.0;
    *rho_m = rm;
    *rho_c = rc;
    *red_grav = rg;
    if (sp_heat_flag) eval_sp_heat(o, alphas, inv_rhom, cp_c, cp_mix);
}

/* eval_local_speeds_x / _y: src/constitutive_2d.f90:2203-2307 */
static void eval_local_speeds(const oracle_t *o, const double *qp, double G, int dir,
                              double *vmin, double *vmax)
{
    double Ri, rm, rc, rg, c1, c2;
    mixt_var(o, qp, &Ri, &rm, &rc, &rg, 0, &c1, &c2);
    double h = qp[0];
    double w = (dir == 1) ? qp[o->nv] : qp[o->nv + 1];
    if (rg * h < 0.0) {
        *vmin = w;
        *vmax = w;
    } else {
        double cel = (dir == 1) ? sqrt(rg * h * G) : sqrt(G * rg * h);
        *vmin = w - cel;
        *vmax = w + cel;
    }
}

/* eval_fluxes: src/constitutive_2d.f90:2325-2518 (:2370-2421 x, :2433-2483 y) */
static void eval_fluxes(const oracle_t *o, const double *qc, const double *qp, double G,
                        int dir, double *f)
{
    const int ns = o->ns, nv = o->nv;
    if (qp[0] > EPS) {
        double h = qp[0];
        double w = (dir == 1) ? qp[nv] : qp[nv + 1];
        double Ri, rm, rc, rg, c1, c2;
        mixt_var(o, qp, &Ri, &rm, &rc, &rg, 0, &c1, &c2);
        double P = 0.5 * rm * G * rg * (h * h);
        f[0] = w * qc[0] * 1.0;
        if (dir == 1) {
            f[1] = w * qc[1] * 1.0 + P;
            f[2] = w * qc[2] * 1.0;
        } else {
            f[1] = w * qc[1] * 1.0;
            f[2] = w * qc[2] * 1.0 + P;
        }
        if (o->energy)
            f[3] = w * (qc[3] * 1.0 + P);
        else
            f[3] = w * qc[3] * 1.0;
        for (int i = 0; i < ns; i++) f[4 + i] = w * qc[4 + i] * 1.0;
        double S = 0.0;
        for (int i = 0; i < ns; i++) S += f[4 + i];
        if ((f[0] > 0.0) && (S / f[0] > 1.0)) {
            for (int i = 0; i < ns; i++) f[4 + i] = f[4 + i] / S * f[0];
        }
        /* fluxes of q1 * Z and q1 * p_exc: src/constitutive_2d.f90:2428-2439,
         * :2490-2501 */
        if (o->stoch_tr) f[o->ist] = w * qc[o->ist] * 1.0;
        if (o->pore) f[o->ipp] = w * qc[o->ipp] * 1.0;
    } else {
        for (int i = 0; i < nv; i++) f[i] = 0.0;
    }
}

/* eval_expl_terms (gravity + lateral radial volume source):
 * src/constitutive_2d.f90:3024-3431 (:3096-3145 gravity, :3317-3427 source) */
static void eval_expl_terms(const oracle_t *o, double Bx, double By, double G, double dGdx,
                            double dGdy, const double *qp, double *ex, double time,
                            double cell_fract, double arc_p, double arc_x, double arc_y,
                            double cell_area)
{
    const int ns = o->ns, nv = o->nv;
    for (int i = 0; i < nv; i++) ex[i] = 0.0;
    if ((qp[0] <= EPS) && (cell_fract == 0.0) && (arc_p == 0.0)) return;
    if (qp[0] > EPS) {
        double h = qp[0], u = qp[nv], v = qp[nv + 1];
        double Ri, rm, rc, rg, cpc, cpm;
        mixt_var(o, qp, &Ri, &rm, &rc, &rg, 1, &cpc, &cpm);
        double centr = 0.0;
        double tg = rg + centr;
        ex[1] = -(G * rm * tg * h * Bx) + 0.5 * rm * rg * (h * h) * dGdx;
        ex[2] = -(G * rm * tg * h * By) + 0.5 * rm * rg * (h * h) * dGdy;
        if (o->energy)
            ex[3] = ex[1] * u + ex[2] * v;
        else
            ex[3] = 0.0;
    }
    if (o->radial && (arc_p > 0.0) && (o->h_source > 0.0)) {
        const double *tp = o->time_param;
        double t_coeff;
        if (time >= tp[3]) {
            if ((tp[2] > 0.0) && (time < tp[3] + tp[2]))
                t_coeff = 0.5 * (1.0 + cos(o->pi_g * (time - tp[3]) / tp[2]));
            else
                t_coeff = 0.0;
        } else {
            double t_rem = fmod(time, tp[0]);
            t_coeff = 0.0;
            if (tp[2] == 0.0) {
                if (t_rem <= tp[1]) t_coeff = 1.0;
            } else {
                if (t_rem < tp[2])
                    t_coeff = 0.5 * (1.0 - cos(o->pi_g * t_rem / tp[2]));
                else if (t_rem <= (tp[1] - tp[2]))
                    t_coeff = 1.0;
                else if (t_rem <= tp[1])
                    t_coeff = 0.5 * (1.0 + cos(o->pi_g * ((t_rem - tp[1]) / tp[2] + 1.0)));
            }
        }
        if (t_coeff > 0.0) {
            double qs[10];  /* n_vars + 2 <= 9 */
            for (int i = 0; i < nv + 2; i++) qs[i] = 0.0;
            qs[0] = 1.0;
            qs[3] = o->T_source;
            for (int i = 0; i < ns; i++) qs[4 + i] = o->alphas_source[i] * qs[0];
            qs[nv] = 0.0;
            qs[nv + 1] = 0.0;
            double Ri, rm, rc, rg, cpc, cpm;
            mixt_var(o, qs, &Ri, &rm, &rc, &rg, 1, &cpc, &cpm);
            double h_dot = o->h_source * o->vel_source * arc_p / cell_area;
            ex[0] = ex[0] + t_coeff * h_dot * rm;
            ex[1] = ex[1] + t_coeff * h_dot * rm * o->vel_source * arc_x;
            ex[2] = ex[2] + t_coeff * h_dot * rm * o->vel_source * arc_y;
            ex[3] = ex[3] + t_coeff * h_dot * rm * cpm * o->T_source;
            for (int i = 0; i < ns; i++)
                ex[4 + i] = ex[4 + i] + t_coeff * h_dot * o->alphas_source[i] * o->rho_s[i];
        }
    }
}

/* sauter_diameter and average_density_solids: src/constitutive_2d.f90:387-416 */
static double sauter_diameter(const oracle_t *o, const double *al)
{
    double s = 0.0, t = 0.0;
    for (int i = 0; i < o->ns; i++) s += al[i];
    if (s < EPS) {
        for (int i = 0; i < o->ns; i++) t += o->diam_s[i];
        return t / (double)o->ns;
    }
    for (int i = 0; i < o->ns; i++) t += al[i] / o->diam_s[i];
    return s / t;
}

static double average_density_solids(const oracle_t *o, const double *al)
{
    double s = 0.0, t = 0.0;
    for (int i = 0; i < o->ns; i++) s += al[i];
    if (s < EPS) {
        for (int i = 0; i < o->ns; i++) t += o->rho_s[i];
        return t / (double)o->ns;
    }
    for (int i = 0; i < o->ns; i++) t += al[i] * o->rho_s[i];
    return t / s;
}

/* Carman-Kozeny permeability with the mean sphericity (dynamic
 * permeability): src/constitutive_2d.f90:3860-3876 (complex), :4975-4992
 * (real).  For one class the class's diameter and sphericity; for several
 * the Sauter diameter and the surface-area-weighted sphericity. */
static void perm_grain(const oracle_t *o, const double *al, double *diam, double *sph)
{
    if (o->ns > 1) {
        *diam = sauter_diameter(o, al);
        double a = 0.0, b = 0.0;
        for (int i = 0; i < o->ns; i++) a += al[i] * (o->diam_s[i] * o->diam_s[i]) * o->sph_s[i];
        for (int i = 0; i < o->ns; i++) b += al[i] * (o->diam_s[i] * o->diam_s[i]);
        *sph = a / b;
    } else {
        *diam = o->diam_s[0];
        *sph = o->sph_s[0];
    }
}

/* the smoothstep sum of f_inhibit (STATIC and DYNAMIC modes):
 * 1 - sum_n pascal(n) tt**(N + n + 1), tt = (x - a_t)/(max_pack - a_t)
 * clamped to [0, 1] (src/constitutive_2d.f90:3904-3927, :4911-4934) */
static double finh_sum(const oracle_t *o, double x, double at)
{
    double width = o->max_pack - at;
    double tt = (x - at) / width;
    if (tt < 0.0) tt = 0.0;
    if (tt > 1.0) tt = 1.0;
    double s = 0.0;
    for (int n = 0; n <= o->N_inh; n++) {
        double term = o->pascal[n] * rpow(o, tt, (double)(o->N_inh + n) + 1.0);
        s = s + term;
    }
    return 1.0 - s;
}

/* 10_wp**(-0.28_wp) as the reference build folds it at compile time */
#define TEN_POW_M028 0x1.0cb3903cf06e8p-1

/* the pore-pressure row of eval_implicit_terms (complex):
 * src/constitutive_2d.f90:3844-3993 */
static cx pore_source_c(const oracle_t *o, cx h, cx T, cx rho_m, const cx *alphas, cx exc)
{
    const double h_threshold = 1.0e-10;
    double gamma_gas = o->cp_a / (o->cp_a - o->R_a);
    cx gas_compr = C(1.0 / (gamma_gas * o->pres), 0.0);
    cx rho_gas = cdiv(C(o->pres, 0.0), cmul(C(o->R_a, 0.0), T));
    cx sa = C(0.0, 0.0);
    for (int i = 0; i < o->ns; i++) sa = cadd(sa, alphas[i]);
    cx porosity = csub(C(1.0, 0.0), sa);
    double hyd = o->hyd_perm;
    if (o->dyn_perm) {
        double al[3], diam, sph;
        for (int i = 0; i < o->ns; i++) al[i] = alphas[i].re;
        perm_grain(o, al, &diam, &sph);
        double ds = diam * sph;
        cx num = cmul(cpow_r(o, porosity, 3.0), C(ds * ds, 0.0));
        cx den = cmul(C(150.0, 0.0), cpow_r(o, csub(C(1.0, 0.0), porosity), 2.0));
        hyd = cdiv(num, den).re;
    }
    cx D = cdiv(C(hyd, 0.0),
                cmul(cmul(cmul(porosity, C(o->kin_visc_c, 0.0)), rho_gas), gas_compr));
    if (!((exc.re > 0.0) && (h.re > 0.0))) return C(0.0, 0.0);
    cx f;
    if (o->finh == 0) {
        f = C(1.0, 0.0);
    } else if (o->finh == 1) {
        f = (sa.re < o->max_pack) ? C(1.0, 0.0) : C(0.0, 0.0);
    } else {
        double at = o->alpha_trans;
        if (o->finh == 3) {
            at = cmul(C(TEN_POW_M028, 0.0), cpow_r(o, h, 0.12)).re;
            if (at >= o->max_pack) at = 0.99 * o->max_pack;
        }
        double width = o->max_pack - at;
        double tt = (sa.re - at) / width;
        if (tt < 0.0) tt = 0.0;
        if (tt > 1.0) tt = 1.0;
        cx sc = C(0.0, 0.0);
        for (int n = 0; n <= o->N_inh; n++) {
            double term = o->pascal[n] * rpow(o, tt, (double)(o->N_inh + n) + 1.0);
            sc = cadd(sc, C(term, 0.0));
        }
        f = csub(C(1.0, 0.0), sc);
    }
    double c2 = (o->pi_g / 2.0) * (o->pi_g / 2.0);
    cx hm = (h_threshold > h.re) ? C(h_threshold, 0.0) : h;
    cx r = cmul(cmul(cdiv(cmul(cmul(rho_m, C(c2, 0.0)), D), hm), exc), f);
    return C(-r.re, -r.im);
}

/* eval_nh_semi_impl_terms: src/constitutive_2d.f90:4027-4547; Voellmy
 * Coulomb part (model 1, :4164-4213) and the mu(I) friction of models 11
 * and 12 (:4384-4537, the two branches compute the same terms) */
static void eval_nh_semi_impl_terms(const oracle_t *o, double G, const double *qp, double *si)
{
    const int nv = o->nv;
    for (int i = 0; i < nv; i++) si[i] = 0.0;
    double h = qp[0], u = qp[nv], v = qp[nv + 1];
    double Ri, rm, rc, rg, c1, c2;
    mixt_var(o, qp, &Ri, &rm, &rc, &rg, 0, &c1, &c2);
    double w = 0.0;
    double mod_vel = sqrt(u * u + v * v + w * w);
    if (o->rheology == 1) {
        if (mod_vel > 0.0) {
            double centr = 0.0;
            double tt = o->mu * rm * (rg + centr) * h * sqrt(G);
            tt = fmax_f(0.0, tt);
            /* :4191-4198 */
            if (o->pore) tt = fmax_f(0.0, tt - o->mu * sqrt(G) * qp[o->ipp]);
            tt = fmax_f(0.0, tt);
            si[1] = si[1] - tt * u / mod_vel;
            si[2] = si[2] - tt * v / mod_vel;
        }
    } else if (mod_vel > 0.0) {
        double diam, rhop;
        if (o->ns > 1) {
            double al[3];
            for (int i = 0; i < o->ns; i++) al[i] = qp[4 + i] / qp[0];
            diam = sauter_diameter(o, al);
            rhop = average_density_solids(o, al);
        } else {
            diam = o->diam_s[0];
            rhop = o->rho_s[0];
        }
        if (h < EPS) h = EPS;
        double centr = 0.0;
        double vse = rm * (rg + centr) * h;
        if (o->pore) vse = rm * (rg + centr) * h - qp[o->ipp];  /* :4423-4427, :4502-4506 */
        vse = fmax_f(vse, 0.0);
        double ens = fmax_f(0.0, vse * sqrt(G));
        if (ens < EPS) ens = EPS;
        double shear = 5.0 / 2.0 * mod_vel / (h * sqrt(G));
        double In = diam * shear / sqrt(ens / rhop);  /* inertial number I */
        double mu_I = (o->mu_s * o->I_0 + o->mu_2 * In + o->muI_inf * (In * In)) / (o->I_0 + In);
        double tt = mu_I * ens;
        si[1] = si[1] - tt * (u / mod_vel);
        si[2] = si[2] - tt * (v / mod_vel);
    }
}

/* c_phys_var (gas, no liquid, no extra gas): src/constitutive_2d.f90:1154-1322;
 * eval_implicit_terms, Voellmy turbulent part (models 1 and 12; model 11 has
 * no implicit friction term): :3513-3683, :3994-4009 */
static void eval_implicit_terms_c(const oracle_t *o, const cx *qj, cx *nh)
{
    const int ns = o->ns, nv = o->nv;
    for (int i = 0; i < nv; i++) nh[i] = C(0.0, 0.0);
    cx u, v, rho_m, inv_rhom, T, hh, alphas[3], exc;
    if (qj[0].re > EPS) {
        cx inv_cqj1 = cdiv(C(1.0, 0.0), qj[0]);
        cx xs[3];
        for (int i = 0; i < ns; i++) xs[i] = cmul(qj[4 + i], inv_cqj1);
        exc = o->pore ? cmul(qj[o->ipp], inv_cqj1) : C(0.0, 0.0);  /* :1195 */
        cx xs_tot = C(0.0, 0.0);
        for (int i = 0; i < ns; i++) xs_tot = cadd(xs_tot, xs[i]);
        cx xc = csub(C(1.0, 0.0), xs_tot);
        cx sp_heat_c = cdiv(cadd(cmul(csub(xc, C(0.0, 0.0)), C(o->cp_a, 0.0)), C(0.0, 0.0)), xc);
        cx dot = C(0.0, 0.0);
        for (int i = 0; i < ns; i++) dot = cadd(dot, cmul(cconj(xs[i]), C(o->cp_s[i], 0.0)));
        cx sp_heat_mix = cadd(dot, cmul(xc, sp_heat_c));
        if (qj[0].re > o->eps_sing) {
            if (o->energy) {
                cx ke = cmul(cmul(C(0.5, 0.0), cadd(csq(qj[1]), csq(qj[2]))), inv_cqj1);
                T = cdiv(csub(qj[3], ke), cmul(qj[0], sp_heat_mix));
            } else {
                T = cdiv(qj[3], cmul(qj[0], sp_heat_mix));
            }
            if (T.re <= 0.0) T = C(o->T_amb, 0.0);
        } else {
            T = C(o->T_amb, 0.0);
        }
        cx Rc = cdiv(cadd(cmul(csub(xc, C(0.0, 0.0)), C(o->R_a, 0.0)), C(0.0, 0.0)), xc);
        cx inv_rho_c = cmul(cmul(Rc, T), C(o->inv_pres, 0.0));
        cx d2 = C(0.0, 0.0);
        for (int i = 0; i < ns; i++) d2 = cadd(d2, cmul(cconj(xs[i]), C(o->inv_rho_s[i], 0.0)));
        inv_rhom = cadd(d2, cmul(xc, inv_rho_c));
        rho_m = cdiv(C(1.0, 0.0), inv_rhom);
        /* :1286-1290 */
        for (int i = 0; i < ns; i++) alphas[i] = cmul(cmul(rho_m, xs[i]), C(o->inv_rho_s[i], 0.0));
        hh = cmul(qj[0], inv_rhom);
        if (qj[0].re > o->eps_sing) {
            u = cmul(qj[1], inv_cqj1);
            v = cmul(qj[2], inv_cqj1);
        } else {
            cx c4 = cmul(qj[0], cmul(qj[0], cmul(qj[0], qj[0])));
            cx den = csqrt_c(cadd(c4, C(o->eps_sing4, 0.0)));
            cx s2 = C(1.4142135623730951, 0.0);
            u = cdiv(cmul(cmul(s2, qj[0]), qj[1]), den);
            cx den2 = csqrt_c(cadd(c4, C(o->eps_sing4, 0.0)));
            v = cdiv(cmul(cmul(s2, qj[0]), qj[2]), den2);
        }
    } else {
        u = C(0.0, 0.0);
        v = C(0.0, 0.0);
        rho_m = C(o->rho_a_amb, 0.0);
        inv_rhom = cdiv(C(1.0, 0.0), rho_m);
        T = C(o->T_amb, 0.0);
        hh = C(0.0, 0.0);
        for (int i = 0; i < ns; i++) alphas[i] = C(0.0, 0.0);
        exc = C(0.0, 0.0);
    }
    (void)inv_rhom;
    cx red_grav = cmul(cdiv(csub(rho_m, C(o->rho_a_amb, 0.0)), rho_m), C(o->grav, 0.0));
    cx w = C(0.0, 0.0);
    cx mod_vel2 = cadd(cadd(csq(u), csq(v)), csq(w));
    cx mod_vel = csqrt_c(mod_vel2);
    if ((o->rheology == 1 || o->rheology == 12) && mod_vel.re != 0.0) {
        cx ts = cmul(cdiv(cmul(rho_m, red_grav), C(o->xi_cell, 0.0)), mod_vel2);
        nh[1] = csub(nh[1], cmul(ts, cdiv(u, mod_vel)));
        nh[2] = csub(nh[2], cmul(ts, cdiv(v, mod_vel)));
    }
    if (o->pore) nh[o->ipp] = pore_source_c(o, hh, T, rho_m, alphas, exc);
}

static void eval_implicit_terms_r(const oracle_t *o, const double *q, double *nh)
{
    cx qj[8] = {{0.0, 0.0}}, r[8];
    for (int i = 0; i < o->nv; i++) qj[i] = C(q[i], 0.0);
    eval_implicit_terms_c(o, qj, r);
    for (int i = 0; i < o->nv; i++) nh[i] = r[i].re;
}

/* xi_temp of eval_implicit_terms: xi + Z(j,k) limited to [1, 1e5] with the
 * stochastic friction, xi without (src/constitutive_2d.f90:3663-3673) */
static double cell_xi(const oracle_t *o, size_t c)
{
    if (!o->stochastic) return o->xi;
    double x = o->xi + o->Z[c];
    if (x < 1.0) x = 1.0;
    if (x > 1e5) x = 1e5;
    return x;
}

/* settling_velocity, Schiller-Naumann drag: src/constitutive_2d.f90:5326-5404 */
static double settling_velocity(const oracle_t *o, double diam, double rhos, double rhoc,
                                double ikv)
{
    double inv_sqrt_C_D = 1.0;
    if (rhos <= rhoc) return 0.0;
    double const_part = sqrt(1.3333333333333333 * (rhos / rhoc - 1.0) * diam * o->grav);
    double sv = const_part * inv_sqrt_C_D;
    double Rey = diam * sv * ikv;
    g_wk.settle_calls++;
    if (Rey <= 1000.0) {
        for (int i = 0; i < 20; i++) {
            double old = sv;
            inv_sqrt_C_D = sqrt(Rey / (24.0 * (1.0 + 0.15 * rpow(o, Rey, 0.687))));
            sv = const_part * inv_sqrt_C_D;
            g_wk.settle_its++;
            if (i + 1 > g_wk.settle_max_its) g_wk.settle_max_its = i + 1;
            if (fabs(old - sv) / old < 1.0e-6) return sv;
            Rey = diam * sv * ikv;
        }
    }
    return sv;
}

/* eval_mass_exchange_terms, no pore pressure / vaporisation:
 * src/constitutive_2d.f90:4574-5153 (:4731-4758 erosion, :4718-4870
 * deposition, :5046-5112 terms).  erod: the cell's erodible layer per class;
 * ers receives the erosion rates.  Returns 1 when the reference would pause
 * (negative deposition). */
static int eval_mass_exchange_terms(const oracle_t *o, const double *qp, const double *erod,
                                    double dt, double *dep, double *eq, double *ers, double *topo)
{
    const int ns = o->ns, nv = o->nv;
    *topo = 0.0;
    for (int i = 0; i < ns; i++) { ers[i] = 0.0; dep[i] = 0.0; }
    for (int i = 0; i < nv; i++) eq[i] = 0.0;
    if (qp[0] <= EPS) return 0;
    double h = qp[0], u = qp[nv], v = qp[nv + 1];
    double alphas[3];
    for (int i = 0; i < ns; i++) alphas[i] = qp[4 + i] / qp[0];
    double at = 0.0;
    for (int i = 0; i < ns; i++) at += alphas[i];
    double w = 0.0;
    double mod_vel2 = u * u + v * v + w * w;
    double mod_vel = sqrt(mod_vel2);
    if (o->erosion_coeff > 0.0) {
        double tot_erosion = o->erosion_coeff * mod_vel * h * (1.0 - at);
        double tse = tot_erosion * (1.0 - o->erodible_porosity);
        for (int i = 0; i < ns; i++) ers[i] = o->erodible_fract[i] * tse;
    }
    for (int i = 0; i < ns; i++) ers[i] = fmax_f(0.0, fmin_f(ers[i], erod[i] / dt));
    double tse = 0.0;
    for (int i = 0; i < ns; i++) tse += ers[i];
    double tot_erosion = tse / (1.0 - o->erodible_porosity);
    double cpe = tot_erosion * o->erodible_porosity;
    if (at <= 0.0) return 0; /* alphastot_min = 0 */
    double T = qp[3];
    double Ri, rm, rc, rg, cpc, cpm;
    mixt_var(o, qp, &Ri, &rm, &rc, &rg, 1, &cpc, &cpm);
    double ikv = 1.0 / o->kin_visc_c;
    double rhoc = rc;
    int err = 0;
    for (int i = 0; i < ns; i++) {
        if ((alphas[i] > 0.0) && o->settling) {
            double sv = settling_velocity(o, o->diam_s[i], o->rho_s[i], rhoc, ikv);
            dep[i] = alphas[i] * sv * 1.0;
            dep[i] = dep[i] * rpow(o, 1.0 - fmin_f(1.0, at / 0.6), 4.65);
            dep[i] = fmin_f(dep[i], h * alphas[i] / dt);
            if (dep[i] < 0.0) err = 1;
        }
    }
    double sdep = 0.0;
    for (int i = 0; i < ns; i++) sdep += dep[i];
    double cpl = 0.0;
    cpl = cpl + o->coeff_porosity * sdep;
    /* gas loss driven by the excess pore pressure, inhibited by f_inhibit:
     * src/constitutive_2d.f90:4887-5032 */
    if (o->pore && o->gas_loss && at < o->max_pack) {
        const double pexc = qp[o->ipp];
        double f_inh = 1.0, hl = h;
        if (o->finh == 1) {
            f_inh = (at < o->max_pack) ? 1.0 : 0.0;
        } else if (o->finh == 2) {
            f_inh = finh_sum(o, at, o->alpha_trans);
        } else if (o->finh == 3) {
            if (hl < EPS) hl = EPS;
            double atd = TEN_POW_M028 * rpow(o, hl, 0.12);
            if (atd >= o->max_pack) atd = 0.99 * o->max_pack;
            f_inh = finh_sum(o, at, atd);
        }
        double hyd = o->hyd_perm;
        if (o->dyn_perm) {
            double diam, sph;
            perm_grain(o, alphas, &diam, &sph);
            double ds = diam * sph, pr = 1.0 - at;
            hyd = ((pr * pr) * pr * (ds * ds)) / (150.0 * (at * at));
        }
        /* MAX(1.e-5, r_h): a default-real literal */
        const double hmin = (double)1.0e-5f;
        double vel = hyd / (o->kin_visc_c * rc) / fmax_f(hmin, hl) * 0.5 * o->pi_g * pexc;
        cpl = cpl + vel * f_inh;
        h = hl;
    }
    /* limits by the maximum solid packing (src/inpout_2d.f90:854) and
     * by the continuous phase present: src/constitutive_2d.f90:5034-5041 */
    cpl = fmin_f(cpl, h * fmax_f(0.0, o->max_pack - at) / dt);
    cpl = fmin_f(cpl, (h * (1.0 - at) - o->coeff_porosity * (h * at - dt * sdep)) / dt);
    const double T_erodible = o->erodible_deposit ? T : o->T_erodible;
    double air = 0.0;
    if (o->entrainment && (h > 0.0) && (Ri > 0.0)) {
        double ec = 0.075 / sqrt(1.0 + 718.0 * rpow(o, Ri, 2.4));
        air = ec * sqrt(mod_vel2);
    }
    double rdt = 0.0, rert = 0.0;
    for (int i = 0; i < ns; i++) rdt += o->rho_s[i] * dep[i];
    for (int i = 0; i < ns; i++) rert += o->rho_s[i] * ers[i];
    eq[0] = o->rho_a_amb * air + rert - rdt + o->rho_c_sub * cpe - rc * cpl;
    eq[1] = -(u * (rdt + rc * cpl));
    eq[2] = -(v * (rdt + rc * cpl));
    double A = 0.0, D = 0.0;
    for (int i = 0; i < ns; i++) A += o->rho_s[i] * o->cp_s[i] * dep[i];
    for (int i = 0; i < ns; i++) D += o->rho_s[i] * o->cp_s[i] * ers[i];
    if (o->energy)
        eq[3] = -(T * (A + rc * cpc * cpl)) - 0.5 * mod_vel2 * (rdt + rc * cpl) +
                T_erodible * (D + o->rho_c_sub * cpc * cpe) +
                o->T_amb * o->cp_a * o->rho_a_amb * air;
    else
        eq[3] = -(T * (A + rc * cpc * cpl)) + T_erodible * (D + o->rho_c_sub * cpc * cpe) +
                o->T_amb * o->cp_a * o->rho_a_amb * air;
    for (int i = 0; i < ns; i++) eq[4 + i] = o->rho_s[i] * (ers[i] - dep[i]);
    if (o->stoch_tr) eq[o->ist] = eq[0] * qp[o->ist];  /* :5132-5137 */
    if (o->pore) eq[o->ipp] = eq[0] * qp[o->ipp];  /* :5140-5145 */
    /* topo_term: src/constitutive_2d.f90:5149 */
    double dep_tot = 0.0, ers_tot = 0.0;
    for (int i = 0; i < ns; i++) dep_tot += dep[i];
    for (int i = 0; i < ns; i++) ers_tot += ers[i];
    *topo = (dep_tot - ers_tot) / (1.0 - o->erodible_porosity);
    return err;
}

/* ====================================================================== */
/* solver_2d                                                               */
/* ====================================================================== */

/* limit, minmod only: src/geometry_2d.f90:1776-1806, :1855-1876 */
static double limit_minmod(const double *vv, const double *z)
{
    double a = (vv[2] - vv[1]) / (z[2] - z[1]);
    double b = (vv[1] - vv[0]) / (z[1] - z[0]);
    double m = fmin_f(fabs(a), fabs(b));
    if (m <= 0.0) return 0.0;
    double sa = a / fabs(a), sb = b / fabs(b);
    return 0.5 * (sa + sb) * m;
}

/* check_solve: src/solver_2d.f90:667-858 */
static void check_solve(oracle_t *o, int solve_all)
{
    const int nx = o->nx, ny = o->ny;
    for (int k = 1; k < ny - 1; k++)
        for (int j = 1; j < nx - 1; j++) o->mask[CELL(o, j, k)] = solve_all ? 1 : 0;
    for (int k = 1; k < ny - 1; k++)
        for (int j = 1; j < nx - 1; j++) {
            int c = CELL(o, j, k);
            if ((o->q[o->nv * c] > 0.0) && (0.0 <= o->t)) o->mask[c] = 1;
        }
    if (o->radial)
        for (int k = 1; k < ny - 1; k++)
            for (int j = 1; j < nx - 1; j++)
                if (o->scell[CELL(o, j, k)] == 2) o->mask[CELL(o, j, k)] = 1;
    for (int it = 0; it < 2; it++) {
        memcpy(o->mtmp, o->mask, (size_t)nx * ny);
        for (int k = 1; k < ny - 1; k++)
            for (int j = 1; j < nx - 1; j++) {
                int c = CELL(o, j, k);
                o->mask[c] = o->mask[c] | o->mtmp[c - 1] | o->mtmp[c + 1] | o->mtmp[c - nx] |
                             o->mtmp[c + nx];
            }
    }
    if (o->radial)
        for (int c = 0; c < nx * ny; c++)
            if (o->scell[c] == 1) o->mask[c] = 0;
    memset(o->mx, 0, (size_t)(nx + 1) * ny);
    memset(o->my, 0, (size_t)nx * (ny + 1));
    int n = 0;
    for (int k = 0; k < ny; k++)
        for (int j = 0; j < nx; j++)
            if (o->mask[CELL(o, j, k)]) {
                o->jc[n] = j;
                o->kc[n] = k;
                n++;
                o->mx[FX(o, j, k)] = 1;
                o->mx[FX(o, j + 1, k)] = 1;
                o->my[FY(o, j, k)] = 1;
                o->my[FY(o, j, k + 1)] = 1;
            }
    o->ncells = n;
    n = 0;
    for (int k = 0; k < ny; k++)
        for (int j = 0; j <= nx; j++)
            if (o->mx[FX(o, j, k)]) { o->jsx[n] = j; o->ksx[n] = k; n++; }
    o->nfx = n;
    n = 0;
    for (int k = 0; k <= ny; k++)
        for (int j = 0; j < nx; j++)
            if (o->my[FY(o, j, k)]) { o->jsy[n] = j; o->ksy[n] = k; n++; }
    o->nfy = n;
}

static void qp_to_qp2(const double *qp, double *qp2)
{
    /* src/constitutive_2d.f90:2164-2188 (qp2(1) = h + B is not used here) */
    if (qp[0] <= 0.0) {
        qp2[1] = 0.0;
        qp2[2] = 0.0;
    } else {
        qp2[1] = qp[1] / qp[0];
        qp2[2] = qp[2] / qp[0];
    }
}

/* stale-state instrumentation: a neighbour's primitive state that was not
 * written for this reconstruction call and is not the dry state a fresh
 * evaluation of a dry cell gives */
static inline void note_stale_qp(oracle_t *o, int c, const double *qpx, int reader)
{
    if (o->stQP[c] != o->call_id && qpx[(size_t)(o->nv + 2) * c] != 0.0) {
        o->stale++;
        o->stale_cell[c] = 1;
        o->stale_cell[reader] = 1;
    }
}

/* reconstruction: src/solver_2d.f90:3095-3924 */
/* the primitive state a fresh evaluation gives a dry cell outside the
 * source disc: zeros, T = T_ambient */
static inline void dry_qp(const oracle_t *o, double *qp)
{
    for (int i = 0; i < o->nv + 2; i++) qp[i] = 0.0;
    qp[3] = o->T_amb;
}

/* Fresh mode: the primitive states of the cells next to the active set that
 * this call does not evaluate (dry cells outside the set, whose stored state
 * may be left from a step when they were wet) are set to what a fresh
 * evaluation gives them, before the active cells' reconstruction reads them. */
static void fresh_neighbour_qp(oracle_t *o, double *qp_expl)
{
    const int nx = o->nx, ny = o->ny, np = o->nv + 2;
    for (int l = 0; l < o->ncells; l++) {
        const int j = o->jc[l], k = o->kc[l];
        const int c = CELL(o, j, k);
        const int nb[4] = {j > 0 ? c - 1 : -1, j < nx - 1 ? c + 1 : -1, k > 0 ? c - nx : -1,
                           k < ny - 1 ? c + nx : -1};
        for (int m = 0; m < 4; m++)
            if (nb[m] >= 0 && o->stQP[nb[m]] != o->call_id && o->scell[nb[m]] != 1)
                dry_qp(o, qp_expl + (size_t)np * nb[m]);
    }
}

static void reconstruction(oracle_t *o, const double *q_expl, const double *qp_expl)
{
    const int nx = o->nx, ny = o->ny, nv = o->nv, np = nv + 2;
    double qrecW[10], qrecE[10], qrecS[10], qrecN[10], sbd[10];  /* n_vars + 2 <= 9 */
    double pr_x[8], pr_y[8], st[3], xs[3], ys[3], q2[3];
    (void)q_expl;
    o->call_id++;
    for (int l = 0; l < o->ncells; l++) o->stQP[CELL(o, o->jc[l], o->kc[l])] = o->call_id;
    if (o->fresh) fresh_neighbour_qp(o, (double *)qp_expl);
    for (int l = 0; l < o->ncells; l++) {
        const int j = o->jc[l], k = o->kc[l];
        const int c = CELL(o, j, k);
        const double *qpc = qp_expl + (size_t)np * c;
        const int src2 = (o->scell[c] == 2);
        for (int i = 0; i < np; i++) {
            qrecW[i] = qpc[i]; qrecE[i] = qpc[i]; qrecS[i] = qpc[i]; qrecN[i] = qpc[i];
            sbd[i] = qpc[i];
        }
        xs[1] = o->x_comp[j];
        ys[1] = o->y_comp[k];
        if (j > 0) note_stale_qp(o, c - 1, qp_expl, c);
        if (j < nx - 1) note_stale_qp(o, c + 1, qp_expl, c);
        if (k > 0) note_stale_qp(o, c - nx, qp_expl, c);
        if (k < ny - 1) note_stale_qp(o, c + nx, qp_expl, c);
        for (int i = 0; i < nv; i++) {
            st[1] = qpc[i];
            /* x direction */
            if (j == 0) {
                xs[0] = o->x_stag[0];
                xs[2] = o->x_comp[j + 1];
                if (src2) {
                    st[0] = sbd[i];
                    st[2] = qp_expl[(size_t)np * (c + 1) + i];
                    pr_x[i] = limit_minmod(st, xs);
                } else if (o->bcf[0][i] == 0) {
                    st[0] = o->bcv[0][i];
                    st[2] = qp_expl[(size_t)np * (c + 1) + i];
                    pr_x[i] = limit_minmod(st, xs);
                } else if (o->bcf[0][i] == 1) {
                    pr_x[i] = o->bcv[0][i];
                } else {
                    pr_x[i] = (qp_expl[(size_t)np * (c + 1) + i] - qp_expl[(size_t)np * c + i]) *
                              o->one_by_dx;
                }
            } else if (j == nx - 1) {
                xs[2] = o->x_stag[nx];
                xs[0] = o->x_comp[j - 1];
                if (src2) {
                    st[2] = sbd[i];
                    st[0] = qp_expl[(size_t)np * (c - 1) + i];
                    pr_x[i] = limit_minmod(st, xs);
                } else if (o->bcf[1][i] == 0) {
                    st[2] = o->bcv[1][i];
                    st[0] = qp_expl[(size_t)np * (c - 1) + i];
                    pr_x[i] = limit_minmod(st, xs);
                } else if (o->bcf[1][i] == 1) {
                    pr_x[i] = o->bcv[1][i];
                } else {
                    pr_x[i] = (qp_expl[(size_t)np * c + i] - qp_expl[(size_t)np * (c - 1) + i]) *
                              o->one_by_dx;
                }
            } else {
                xs[0] = o->x_comp[j - 1];
                xs[2] = o->x_comp[j + 1];
                st[0] = qp_expl[(size_t)np * (c - 1) + i];
                st[2] = qp_expl[(size_t)np * (c + 1) + i];
                if (o->radial && src2) {
                    if (o->sE[c]) {
                        xs[2] = o->x_stag[j + 1];
                        st[2] = sbd[i];
                    } else if (o->sW[c]) {
                        xs[0] = o->x_stag[j];
                        st[0] = sbd[i];
                    }
                }
                pr_x[i] = limit_minmod(st, xs);
            }
            double dq = 1.0 * o->dx2 * pr_x[i];
            qrecW[i] = st[1] - dq;
            qrecE[i] = st[1] + dq;
            if (j == 0) {
                if (src2) qrecW[i] = sbd[i];
                else if (o->bcf[0][i] == 0) qrecW[i] = o->bcv[0][i];
                else if (i == 1) qrecW[i] = fmin_f(qrecW[i], 0.0);
            }
            if (j == nx - 1) {
                if (src2) qrecE[i] = sbd[i];
                else if (o->bcf[1][i] == 0) qrecE[i] = o->bcv[1][i];
                else if (i == 1) qrecE[i] = fmax_f(qrecE[i], 0.0);
            }
            /* y direction */
            if (k == 0) {
                ys[0] = o->y_stag[0];
                ys[2] = o->y_comp[k + 1];
                if (o->bcf[2][i] == 0) {
                    st[0] = o->bcv[2][i];
                    st[2] = qp_expl[(size_t)np * (c + nx) + i];
                    pr_y[i] = limit_minmod(st, ys);
                } else if (o->bcf[2][i] == 1) {
                    pr_y[i] = o->bcv[2][i];
                } else {
                    pr_y[i] = (qp_expl[(size_t)np * (c + nx) + i] - qp_expl[(size_t)np * c + i]) *
                              o->one_by_dy;
                }
            } else if (k == ny - 1) {
                ys[0] = o->y_comp[k - 1];
                ys[2] = o->y_stag[ny];
                if (o->bcf[3][i] == 0) {
                    st[2] = o->bcv[3][i];
                    st[0] = qp_expl[(size_t)np * (c - nx) + i];
                    pr_y[i] = limit_minmod(st, ys);
                } else if (o->bcf[3][i] == 1) {
                    pr_y[i] = o->bcv[3][i];
                } else {
                    pr_y[i] = (qp_expl[(size_t)np * c + i] - qp_expl[(size_t)np * (c - nx) + i]) *
                              o->one_by_dy;
                }
            } else {
                ys[0] = o->y_comp[k - 1];
                ys[2] = o->y_comp[k + 1];
                st[0] = qp_expl[(size_t)np * (c - nx) + i];
                st[2] = qp_expl[(size_t)np * (c + nx) + i];
                if (o->radial && src2) {
                    if (o->sS[c]) {
                        ys[0] = o->y_stag[k];
                        st[0] = sbd[i];
                    } else if (o->sN[c]) {
                        ys[2] = o->y_stag[k + 1];
                        st[2] = sbd[i];
                    }
                }
                pr_y[i] = limit_minmod(st, ys);
            }
            dq = 1.0 * o->dy2 * pr_y[i];
            qrecS[i] = st[1] - dq;
            qrecN[i] = st[1] + dq;
            if (k == 0) {
                if (o->bcf[2][i] == 0) qrecS[i] = o->bcv[2][i];
                else if (i == 2) qrecS[i] = fmin_f(qrecS[i], 0.0);
            }
            if (k == ny - 1) {
                if (o->bcf[3][i] == 0) qrecN[i] = o->bcv[3][i];
                else if (i == 2) qrecN[i] = fmax_f(qrecN[i], 0.0);
            }
        }
        for (int i = nv; i < nv + 2; i++) {
            /* x direction, u and v with the limiters of hu, hv */
            st[1] = qpc[i];
            if (j == 0) {
                qp_to_qp2(qrecW, q2);
                st[0] = q2[i - nv + 1];
                st[2] = qp_expl[(size_t)np * (c + 1) + i];
            } else if (j == nx - 1) {
                qp_to_qp2(qrecE, q2);
                st[0] = qp_expl[(size_t)np * (c - 1) + i];
                st[2] = q2[i - nv + 1];
            } else {
                st[0] = qp_expl[(size_t)np * (c - 1) + i];
                st[2] = qp_expl[(size_t)np * (c + 1) + i];
                if (o->radial && src2) {
                    if (o->sE[c]) {
                        xs[2] = o->x_stag[j + 1];
                        st[2] = sbd[i];
                    } else if (o->sW[c]) {
                        xs[0] = o->x_stag[j];
                        st[0] = sbd[i];
                    }
                }
            }
            pr_x[i] = limit_minmod(st, xs);
            double dq = 1.0 * o->dx2 * pr_x[i];
            qrecW[i] = st[1] - dq;
            qrecE[i] = st[1] + dq;
            if (j == 0) {
                qp_to_qp2(qrecW, q2);
                qrecW[i] = q2[i - nv + 1];
            } else if (j == nx - 1) {
                qp_to_qp2(qrecE, q2);
                qrecE[i] = q2[i - nv + 1];
            } else if (o->radial && src2) {
                if (o->sE[c]) {
                    for (int m = 0; m < np; m++) qrecE[m] = sbd[m];
                } else if (o->sW[c]) {
                    for (int m = 0; m < np; m++) qrecW[m] = sbd[m];
                }
            }
            /* y direction */
            st[1] = qpc[i];
            if (k == 0) {
                qp_to_qp2(qrecS, q2);
                st[0] = q2[i - nv + 1];
                st[2] = qp_expl[(size_t)np * (c + nx) + i];
            } else if (k == ny - 1) {
                qp_to_qp2(qrecN, q2);
                st[0] = qp_expl[(size_t)np * (c - nx) + i];
                st[2] = q2[i - nv + 1];
            } else {
                st[0] = qp_expl[(size_t)np * (c - nx) + i];
                st[2] = qp_expl[(size_t)np * (c + nx) + i];
                if (o->radial && src2) {
                    if (o->sS[c]) {
                        ys[0] = o->y_stag[k];
                        st[0] = sbd[i];
                    } else if (o->sN[c]) {
                        /* the reference stores this stencil point in the x
                         * stencil (src/solver_2d.f90:3607) */
                        xs[2] = o->y_stag[k + 1];
                        st[2] = sbd[i];
                    }
                }
            }
            pr_y[i] = limit_minmod(st, ys);
            dq = 1.0 * o->dy2 * pr_y[i];
            qrecS[i] = st[1] - dq;
            qrecN[i] = st[1] + dq;
            if (k == 0) {
                qp_to_qp2(qrecS, q2);
                qrecS[i] = q2[i - nv + 1];
            } else if (k == ny - 1) {
                qp_to_qp2(qrecN, q2);
                qrecN[i] = q2[i - nv + 1];
            } else if (o->radial && src2) {
                if (o->sS[c]) {
                    for (int m = 0; m < np; m++) qrecS[m] = sbd[m];
                } else if (o->sN[c]) {
                    for (int m = 0; m < np; m++) qrecN[m] = sbd[m];
                }
            }
        }
        /* x faces */
        {
            const int fw = FX(o, j, k), fe = FX(o, j + 1, k);
            qp_to_qc(o, qrecW, o->qR + (size_t)nv * fw);
            qp_to_qc(o, qrecE, o->qL + (size_t)nv * fe);
            memcpy(o->qpR + (size_t)np * fw, qrecW, sizeof(double) * np);
            memcpy(o->qpL + (size_t)np * fe, qrecE, sizeof(double) * np);
            o->stR[fw] = o->call_id;
            o->stL[fe] = o->call_id;
            if (j == 0) o->stL[fw] = o->call_id;
            if (j == nx - 1) o->stR[fe] = o->call_id;
            if (o->radial && src2 && o->sE[c]) o->stR[fe] = o->call_id;
            if (o->radial && src2 && o->sW[c]) o->stL[fw] = o->call_id;
            if (j == 0) {
                memcpy(o->qL + (size_t)nv * fw, o->qR + (size_t)nv * fw, sizeof(double) * nv);
                memcpy(o->qpL + (size_t)np * fw, o->qpR + (size_t)np * fw, sizeof(double) * np);
            } else if (j == nx - 1) {
                memcpy(o->qR + (size_t)nv * fe, o->qL + (size_t)nv * fe, sizeof(double) * nv);
                memcpy(o->qpR + (size_t)np * fe, o->qpL + (size_t)np * fe, sizeof(double) * np);
            } else if (o->radial && src2) {
                if (o->sE[c]) {
                    memcpy(o->qR + (size_t)nv * fe, o->qL + (size_t)nv * fe, sizeof(double) * nv);
                    o->qR[(size_t)nv * fe + 1] = -o->qL[(size_t)nv * fe + 1];
                    memcpy(o->qpR + (size_t)np * fe, o->qpL + (size_t)np * fe, sizeof(double) * np);
                    o->qpR[(size_t)np * fe + nv] = -o->qpL[(size_t)np * fe + nv];
                } else if (o->sW[c]) {
                    memcpy(o->qL + (size_t)nv * fw, o->qR + (size_t)nv * fw, sizeof(double) * nv);
                    o->qL[(size_t)nv * fw + 1] = -o->qR[(size_t)nv * fw + 1];
                    memcpy(o->qpL + (size_t)np * fw, o->qpR + (size_t)np * fw, sizeof(double) * np);
                    o->qpL[(size_t)np * fw + nv] = -o->qpR[(size_t)np * fw + nv];
                }
            }
        }
        /* y faces */
        {
            const int fs = FY(o, j, k), fn = FY(o, j, k + 1);
            qp_to_qc(o, qrecS, o->qT + (size_t)nv * fs);
            qp_to_qc(o, qrecN, o->qB + (size_t)nv * fn);
            memcpy(o->qpT + (size_t)np * fs, qrecS, sizeof(double) * np);
            memcpy(o->qpB + (size_t)np * fn, qrecN, sizeof(double) * np);
            o->stT[fs] = o->call_id;
            o->stB[fn] = o->call_id;
            if (k == 0) o->stB[fs] = o->call_id;
            if (k == ny - 1) o->stT[fn] = o->call_id;
            if (o->radial && src2 && o->sS[c]) o->stB[fs] = o->call_id;
            if (o->radial && src2 && o->sN[c]) o->stT[fn] = o->call_id;
            if (k == 0) {
                memcpy(o->qB + (size_t)nv * fs, o->qT + (size_t)nv * fs, sizeof(double) * nv);
                memcpy(o->qpB + (size_t)np * fs, o->qpT + (size_t)np * fs, sizeof(double) * np);
            } else if (k == ny - 1) {
                memcpy(o->qT + (size_t)nv * fn, o->qB + (size_t)nv * fn, sizeof(double) * nv);
                memcpy(o->qpT + (size_t)np * fn, o->qpB + (size_t)np * fn, sizeof(double) * np);
            } else if (o->radial && src2) {
                if (o->sS[c]) {
                    memcpy(o->qB + (size_t)nv * fs, o->qT + (size_t)nv * fs, sizeof(double) * nv);
                    o->qB[(size_t)nv * fs + 2] = -o->qT[(size_t)nv * fs + 2];
                    memcpy(o->qpB + (size_t)np * fs, o->qpT + (size_t)np * fs, sizeof(double) * np);
                    o->qpB[(size_t)np * fs + nv + 1] = -o->qpT[(size_t)np * fs + nv + 1];
                } else if (o->sN[c]) {
                    memcpy(o->qT + (size_t)nv * fn, o->qB + (size_t)nv * fn, sizeof(double) * nv);
                    o->qT[(size_t)nv * fn + 2] = -o->qB[(size_t)nv * fn + 2];
                    memcpy(o->qpT + (size_t)np * fn, o->qpB + (size_t)np * fn, sizeof(double) * np);
                    o->qpT[(size_t)np * fn + nv + 1] = -o->qpB[(size_t)np * fn + nv + 1];
                }
            }
        }
    }
}

/* a face side not written by this reconstruction call that is not dry */
static inline void note_stale_face(oracle_t *o, int stamp, const double *qpf, int ca, int cb)
{
    if (stamp != o->call_id && qpf[0] != 0.0) {
        o->stale++;
        if (ca >= 0) o->stale_cell[ca] = 1;
        if (cb >= 0) o->stale_cell[cb] = 1;
    }
}

/* eval_speeds: src/solver_2d.f90:3937-4014 */
/* Fresh mode: a face side this call's reconstruction did not write (the side
 * of a dry cell outside the active set) takes what a fresh evaluation gives
 * it: the dry cell's reconstruction from dry neighbours, the dry primitive
 * state and a zero conservative state. */
static void fresh_face_sides(oracle_t *o)
{
    const int nv = o->nv, np = nv + 2;
    struct { int *st; double *qp, *q; } side[4] = {{o->stL, o->qpL, o->qL}, {o->stR, o->qpR, o->qR},
                                                   {o->stB, o->qpB, o->qB}, {o->stT, o->qpT, o->qT}};
    for (int d = 0; d < 4; d++) {
        const int nf = d < 2 ? o->nfx : o->nfy;
        for (int l = 0; l < nf; l++) {
            const int f = d < 2 ? FX(o, o->jsx[l], o->ksx[l]) : FY(o, o->jsy[l], o->ksy[l]);
            if (side[d].st[f] == o->call_id) continue;
            dry_qp(o, side[d].qp + (size_t)np * f);
            for (int i = 0; i < nv; i++) side[d].q[(size_t)nv * f + i] = 0.0;
        }
    }
}

static void eval_speeds(oracle_t *o)
{
    const int np = o->nv + 2;
    if (o->fresh) fresh_face_sides(o);
    for (int l = 0; l < o->nfx; l++) {
        const int j = o->jsx[l], k = o->ksx[l], f = FX(o, j, k);
        double lmin, lmax, rmin, rmax;
        const int xa = (j > 0) ? CELL(o, j - 1, k) : -1, xb = (j < o->nx) ? CELL(o, j, k) : -1;
        note_stale_face(o, o->stL[f], o->qpL + (size_t)np * f, xa, xb);
        note_stale_face(o, o->stR[f], o->qpR + (size_t)np * f, xa, xb);
        eval_local_speeds(o, o->qpL + (size_t)np * f, 1.0, 1, &lmin, &lmax);
        eval_local_speeds(o, o->qpR + (size_t)np * f, 1.0, 1, &rmin, &rmax);
        o->axn[f] = fmin_f(fmin_f(lmin, rmin), 0.0);
        o->axp[f] = fmax_f(fmax_f(lmax, rmax), 0.0);
    }
    for (int l = 0; l < o->nfy; l++) {
        const int j = o->jsy[l], k = o->ksy[l], f = FY(o, j, k);
        double bmin, bmax, tmin, tmax;
        const int ya = (k > 0) ? CELL(o, j, k - 1) : -1, yb = (k < o->ny) ? CELL(o, j, k) : -1;
        note_stale_face(o, o->stB[f], o->qpB + (size_t)np * f, ya, yb);
        note_stale_face(o, o->stT[f], o->qpT + (size_t)np * f, ya, yb);
        eval_local_speeds(o, o->qpB + (size_t)np * f, 1.0, 2, &bmin, &bmax);
        eval_local_speeds(o, o->qpT + (size_t)np * f, 1.0, 2, &tmin, &tmax);
        o->ayn[f] = fmin_f(fmin_f(bmin, tmin), 0.0);
        o->ayp[f] = fmax_f(fmax_f(bmax, tmax), 0.0);
    }
}

/* eval_flux_KT + average_KT: src/solver_2d.f90:2834-3002, :3019-3049 */
static void eval_flux_KT(oracle_t *o)
{
    const int nv = o->nv, np = nv + 2, ns = o->ns;
    double fL[8], fR[8];
    for (int l = 0; l < o->nfx; l++) {
        const int f = FX(o, o->jsx[l], o->ksx[l]);
        const double *ql = o->qL + (size_t)nv * f, *qr = o->qR + (size_t)nv * f;
        const double *pl = o->qpL + (size_t)np * f, *pr = o->qpR + (size_t)np * f;
        double *H = o->Hx + (size_t)nv * f;
        eval_fluxes(o, ql, pl, 1.0, 1, fL);
        eval_fluxes(o, qr, pr, 1.0, 1, fR);
        const double an = o->axn[f], ap = o->axp[f];
        for (int i = 0; i < nv; i++) {
            if (an == ap) {
                H[i] = 0.0;
            } else {
                double avg = (ap * fL[i] - an * fR[i]) / (ap - an);
                H[i] = avg + (ap * an) / (ap - an) * (qr[i] - ql[i]);
            }
        }
        const int wslot = (o->ksx[l] == o->watch_k) ? o->jsx[l] - o->watch_j : -1;
        if (wslot == 0 || wslot == 1)
            for (int i = 0; i < nv; i++) o->watch_pre[wslot][i] = H[i];
        double S = 0.0;
        for (int i = 0; i < ns; i++) S += H[4 + i];
        if ((S / H[0]) >= 1.0) {
            double S2 = 0.0;
            for (int i = 0; i < ns; i++) S2 += H[4 + i] / H[0];
            for (int i = 0; i < ns; i++) H[4 + i] = H[4 + i] / S2;
        }
        if ((pl[1] == 0.0) && (pr[1] == 0.0)) {
            H[0] = 0.0;
            for (int i = 3; i < nv; i++) H[i] = 0.0;
        }
        if (wslot == 0 || wslot == 1) {
            for (int i = 0; i < nv; i++) o->watch_post[wslot][i] = H[i];
            o->watch_seen[wslot] = 1;
        }
    }
    for (int l = 0; l < o->nfy; l++) {
        const int f = FY(o, o->jsy[l], o->ksy[l]);
        const double *qb = o->qB + (size_t)nv * f, *qt = o->qT + (size_t)nv * f;
        const double *pb = o->qpB + (size_t)np * f, *pt = o->qpT + (size_t)np * f;
        double *H = o->Hy + (size_t)nv * f;
        eval_fluxes(o, qb, pb, 1.0, 2, fL);
        eval_fluxes(o, qt, pt, 1.0, 2, fR);
        const double an = o->ayn[f], ap = o->ayp[f];
        for (int i = 0; i < nv; i++) {
            if (an == ap) {
                H[i] = 0.0;
            } else {
                double avg = (ap * fL[i] - an * fR[i]) / (ap - an);
                H[i] = avg + (ap * an) / (ap - an) * (qt[i] - qb[i]);
            }
        }
        double S = 0.0;
        for (int i = 0; i < ns; i++) S += H[4 + i];
        if ((S / H[0]) > 1.0) {
            double S2 = S / H[0];
            for (int i = 0; i < ns; i++) H[4 + i] = H[4 + i] / S2;
        }
        if ((qb[2] == 0.0) && (qt[2] == 0.0)) {
            H[0] = 0.0;
            for (int i = 3; i < nv; i++) H[i] = 0.0;
        }
    }
}

/* eval_hyperbolic_terms: src/solver_2d.f90:2616-2693 */
static void eval_hyperbolic_terms(oracle_t *o, const double *q_expl, const double *qp_expl,
                                  double *divF)
{
    const int nv = o->nv;
    reconstruction(o, q_expl, qp_expl);
    eval_speeds(o);
    eval_flux_KT(o);
    for (int l = 0; l < o->ncells; l++) {
        const int j = o->jc[l], k = o->kc[l], c = CELL(o, j, k);
        const double *He = o->Hx + (size_t)nv * FX(o, j + 1, k), *Hw = o->Hx + (size_t)nv * FX(o, j, k);
        const double *Hn = o->Hy + (size_t)nv * FY(o, j, k + 1), *Hs = o->Hy + (size_t)nv * FY(o, j, k);
        for (int i = 0; i < nv; i++) {
            double d = 0.0;
            d = d + (He[i] - Hw[i]) * o->one_by_dx;
            d = d + (Hn[i] - Hs[i]) * o->one_by_dy;
            divF[(size_t)nv * c + i] = d;
        }
    }
}

/* timestep: src/solver_2d.f90:873-1006 */
static void timestep(oracle_t *o)
{
    const int nv = o->nv, np = nv + 2;
    o->dt = o->max_dt;
    for (int l = 0; l < o->ncells; l++) {
        const int c = CELL(o, o->jc[l], o->kc[l]);
        double pd;
        if (o->q[(size_t)nv * c] > 0.0) {
            qc_to_qp(o, o->q + (size_t)nv * c, o->qp + (size_t)np * c, &pd);
        } else {
            for (int i = 0; i < np; i++) o->qp[(size_t)np * c + i] = 0.0;
            o->qp[(size_t)np * c + 3] = o->T_amb;
        }
    }
    reconstruction(o, o->q, o->qp);
    eval_speeds(o);
    for (int l = 0; l < o->nfx; l++) {
        const int f = FX(o, o->jsx[l], o->ksx[l]);
        o->axmax[f] = fmax_f(o->axp[f], -o->axn[f]);
    }
    for (int l = 0; l < o->nfy; l++) {
        const int f = FY(o, o->jsy[l], o->ksy[l]);
        o->aymax[f] = fmax_f(o->ayp[f], -o->ayn[f]);
    }
    for (int l = 0; l < o->ncells; l++) {
        const int j = o->jc[l], k = o->kc[l];
        double max_a = fmax_f(o->axmax[FX(o, j, k)], o->axmax[FX(o, j + 1, k)]);
        double dtx = (max_a > 0.0) ? o->cfl * o->dx / max_a : o->dt;
        max_a = fmax_f(o->aymax[FY(o, j, k)], o->aymax[FY(o, j, k + 1)]);
        double dty = (max_a > 0.0) ? o->cfl * o->dy / max_a : o->dt;
        double dtc = fmin_f(dtx, dty);
        o->dt = fmin_f(o->dt, dtc);
    }
}

/* eval_f: src/solver_2d.f90:2245-2284 */
static void eval_f(const oracle_t *o, double dt, const double *qj, const double *qj_old,
                   double a_diag, const double *Rni, double *f, double *scal_f)
{
    const int nv = o->nv;
    double nh[8];
    g_wk.evalf++;
    g_wk.cell_evals++;
    eval_implicit_terms_r(o, qj, nh);
    double s = 0.0;
    for (int i = 0; i < nv; i++) {
        double Rj = Rni[i] - a_diag * nh[i];
        f[i] = qj[i] - qj_old[i] + dt * Rj;
        f[i] = 1.0 * f[i];
    }
    for (int i = 0; i < nv; i++) s += f[i] * f[i];
    *scal_f = 0.5 * s;
}

/* implicit_flag (src/constitutive_2d.f90:478-528): the momentum equations
 * and, with pore pressure, the pore-pressure equation */
static inline int implicit_var(const oracle_t *o, int i)
{
    return i == 1 || i == 2 || (o->pore && i == o->ipp);
}

/* DGESV for the n_nh x n_nh implicit block (n_nh = 3 with pore pressure):
 * src/solver_2d.f90:1905-1917.  The reference links LAPACK from the OpenBLAS
 * build that SciPy bundles; its result is reproduced here by a left-looking
 * (Crout) LU with partial pivoting: column j first takes the earlier row
 * interchanges, then a(i,j) -= dot(a(i,0:m), a(0:m,j)) with the dot product
 * accumulated by fused multiply-adds from 0 (m = i for the rows above the
 * diagonal, m = j below), the pivot is the first largest |a(i,j)|, rows j
 * and p are swapped in columns 0..j, and the column below the pivot is
 * scaled by the reciprocal of the pivot; the solve permutes b, then runs
 * the unit-lower and the upper triangles with fused multiply-adds.  This
 * equals the library's DGESV bit for bit on 10^6 random 3x3 systems
 * (tests/test_oracle_vs_ref.py repeats the check).  A zero pivot (info > 0)
 * leaves b unchanged, as DGESV does.  a: column-major n x n. */
void oracle_dgesv(int n, double *a, double *b)
{
    int ipiv[8], info = 0;
#define A_(i, j) a[(i) + n * (j)]
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++) {
            int p = ipiv[i];
            if (p != i) { double t = A_(i, j); A_(i, j) = A_(p, j); A_(p, j) = t; }
        }
        for (int i = 1; i < j; i++) {
            double d = 0.0;
            for (int k = 0; k < i; k++) d = fma(A_(i, k), A_(k, j), d);
            A_(i, j) -= d;
        }
        for (int i = j; i < n; i++) {
            double d = 0.0;
            for (int k = 0; k < j; k++) d = fma(A_(i, k), A_(k, j), d);
            A_(i, j) -= d;
        }
        int p = j;
        double mx = fabs(A_(j, j));
        for (int i = j + 1; i < n; i++)
            if (fabs(A_(i, j)) > mx) { mx = fabs(A_(i, j)); p = i; }
        ipiv[j] = p;
        if (p != j)
            for (int k = 0; k <= j; k++) { double t = A_(j, k); A_(j, k) = A_(p, k); A_(p, k) = t; }
        if (A_(j, j) != 0.0) {
            double r = 1.0 / A_(j, j);
            for (int i = j + 1; i < n; i++) A_(i, j) *= r;
        } else if (!info) {
            info = j + 1;
        }
    }
    if (info) return;
    for (int j = 0; j < n; j++) {
        int p = ipiv[j];
        if (p != j) { double t = b[j]; b[j] = b[p]; b[p] = t; }
    }
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) b[i] = fma(-b[j], A_(i, j), b[i]);
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= A_(j, j);
        for (int i = 0; i < j; i++) b[i] = fma(-b[j], A_(i, j), b[i]);
    }
#undef A_
}

/* eval_jacobian (complex step): src/solver_2d.f90:2302-2367.  L is
 * column-major nv x nv. */
static void eval_jacobian(const oracle_t *o, double dt, double a_diag, const double *qrel,
                          const double *qorg, double *L)
{
    const int nv = o->nv;
    for (int i = 0; i < nv * nv; i++) L[i] = 0.0;
    cx init[8], qc[8], nh[8];
    g_wk.jac++;
    g_wk.cell_evals += 2;
    for (int i = 0; i < nv; i++) init[i] = C(qrel[i], 0.0);
    for (int i = 0; i < nv; i++) {
        L[i + nv * i] = 1.0 * qorg[i];
        if (implicit_var(o, i)) {
            for (int m = 0; m < nv; m++) qc[m] = init[m];
            qc[i] = C(qrel[i], o->hcs);
            for (int m = 0; m < nv; m++) qc[m] = cmul(qc[m], C(qorg[m], 0.0));
            eval_implicit_terms_c(o, qc, nh);
            for (int r = 0; r < nv; r++) {
                double J = 1.0 * nh[r].im * o->one_by_hcs;
                L[r + nv * i] = L[r + nv * i] - dt * a_diag * J;
            }
        }
    }
}

/* lnsrch: src/solver_2d.f90:2021-2226 */
static void lnsrch(const oracle_t *o, double dt, double a_diag, const double *xold,
                   const double *qorg, const double *qj_old, double fold, const double *g,
                   double *p, double *x, double *f_out, double *rt, double stpmax,
                   const double *Rni)
{
    const int nv = o->nv;
    const double TOLX = EPS, ALF = 1.0e-4;
    (void)ALF;
    double sum = 0.0;
    for (int i = 0; i < nv; i++) sum += p[i] * p[i];
    double pabs = sqrt(sum);
    if (pabs > stpmax)
        for (int i = 0; i < nv; i++) p[i] = p[i] * stpmax / pabs;
    double slope = 0.0;
    for (int i = 0; i < nv; i++) slope += g[i] * p[i];
    double test = 0.0;
    for (int i = 0; i < nv; i++) {
        double tmp = fabs(p[i]) / fmax_f(fabs(xold[i]), 1.0);
        if (i == 0 || tmp > test) test = tmp;
    }
    double alamin = TOLX / test;
    if (alamin == 0.0) {
        for (int i = 0; i < nv; i++) x[i] = xold[i];
        return;
    }
    double alam = 1.0, alam2 = 0.0, f2 = 0.0, tmplam = 0.0, f = 0.0;
    double qj[8];
    for (;;) {
        for (int i = 0; i < nv; i++) x[i] = xold[i] + alam * p[i];
        for (int i = 0; i < nv; i++) qj[i] = x[i] * qorg[i];
        eval_f(o, dt, qj, qj_old, a_diag, Rni, rt, &f);
        g_wk.ls_trials++;
        if (f <= 0.9 * fold) {
            break;
        } else if (alam < alamin) {
            for (int i = 0; i < nv; i++) x[i] = xold[i];
            f = fold;
            break;
        } else {
            if (alam == 1.0) {
                tmplam = -(slope / (2.0 * (f - fold - slope)));
            } else {
                double rhs1 = f - fold - alam * slope;
                double rhs2 = f2 - fold - alam2 * slope;
                double a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
                double b = (-(alam2 * rhs1 / (alam * alam)) + alam * rhs2 / (alam2 * alam2)) /
                           (alam - alam2);
                if (a == 0.0) {
                    tmplam = -(slope / (2.0 * b));
                } else {
                    double disc = b * b - 3.0 * a * slope;
                    if (disc < 0.0)
                        tmplam = 0.5 * alam;
                    else if (b <= 0.0)
                        tmplam = (-b + sqrt(disc)) / (3.0 * a);
                    else
                        tmplam = -(slope / (b + sqrt(disc)));
                }
                if (tmplam > 0.5 * alam) tmplam = 0.5 * alam;
            }
        }
        alam2 = alam;
        f2 = f;
        alam = fmax_f(tmplam, 0.5 * alam);
    }
    *f_out = f;
}

/* solve_rk_step: src/solver_2d.f90:1662-1996 (n_nh = 2 path) */
static int solve_rk_step(const oracle_t *o, double dt, double *qj, const double *qj_old,
                         double a_diag, const double *Rni)
{
    const int nv = o->nv;
    const double TOLF = 1.0e-10, STPMX = 100.0;
    double qorg[8], qrel[8], rt[8], L[64], xold[8], p[8], grad[8];
    double scal_f, scal_f_old;
    for (int i = 0; i < nv; i++) grad[i] = 0.0;
    for (int i = 0; i < nv; i++) qorg[i] = fmax_f(fabs(qj[i]), 1.0e-3);
    for (int i = 0; i < nv; i++) qrel[i] = qj[i] / qorg[i];
    int it;
    for (it = 1; it <= 100; it++) {
        const double TOLX = EPS;
        eval_f(o, dt, qj, qj_old, a_diag, Rni, rt, &scal_f);
        double mx = 0.0;
        for (int i = 0; i < nv; i++) mx = fmax_f(mx, fabs(rt[i]));
        if (mx < TOLF) break;
        eval_jacobian(o, dt, a_diag, qrel, qorg, L);
        /* block solve: explicit rows by their diagonal, the implicit block by
         * Cramer's rule (n_nh = 2) or DGESV (n_nh = 3): :1851-1926 */
        int imp[3], nn = 0;
        for (int r = 0; r < nv; r++)
            if (implicit_var(o, r)) imp[nn++] = r;
        double d1[7], d2[3];
        int ne = 0;
        for (int r = 0; r < nv; r++)
            if (!implicit_var(o, r)) { d1[ne] = rt[r] / L[r + nv * r]; ne++; }
        for (int b = 0; b < nn; b++) d2[b] = rt[imp[b]];
        {
            /* reshape(pack(L, mask21), [nv-nn, nn]) then MATMUL(d1, .) */
            double pk[24];
            int m = 0;
            for (int col = 0; col < nv; col++) {
                if (implicit_var(o, col)) continue;
                for (int b = 0; b < nn; b++) pk[m++] = L[imp[b] + nv * col];
            }
            for (int b = 0; b < nn; b++) {
                double s = 0.0;
                for (int a = 0; a < nv - nn; a++) s += d1[a] * pk[a + (nv - nn) * b];
                d2[b] = d2[b] - s;
            }
        }
        double sol[3];
        if (nn == 2) {
            double l11 = L[1 + nv * 1], l22 = L[2 + nv * 2], l21 = L[2 + nv * 1], l12 = L[1 + nv * 2];
            double inv_det = 1.0 / (l11 * l22 - l21 * l12);
            sol[0] = (d2[0] * l22 - d2[1] * l12) * inv_det;
            sol[1] = (l11 * d2[1] - l21 * d2[0]) * inv_det;
        } else {
            double a22[9];
            for (int c = 0; c < nn; c++)
                for (int r = 0; r < nn; r++) a22[r + nn * c] = L[imp[r] + nv * imp[c]];
            for (int b = 0; b < nn; b++) sol[b] = d2[b];
            oracle_dgesv(nn, a22, sol);
        }
        ne = 0;
        int ni = 0;
        for (int r = 0; r < nv; r++) {
            if (implicit_var(o, r)) {
                p[r] = -sol[ni] + 0.0;
                ni++;
            } else {
                p[r] = 0.0 + -d1[ne];
                ne++;
            }
        }
        for (int i = 0; i < nv; i++) xold[i] = qrel[i];
        scal_f_old = scal_f;
        if (it > 1) {
            double ss = 0.0;
            for (int i = 0; i < nv; i++) ss += qrel[i] * qrel[i];
            double stpmax = STPMX * fmax_f(sqrt(ss), (double)nv);
            for (int i = 0; i < nv; i++) {
                double s = 0.0;
                for (int r = 0; r < nv; r++) s += rt[r] * L[r + nv * i];
                grad[i] = s;
            }
            lnsrch(o, dt, a_diag, xold, qorg, qj_old, scal_f_old, grad, p, qrel, &scal_f, rt,
                   stpmax, Rni);
        } else {
            for (int i = 0; i < nv; i++) qrel[i] = xold[i] + p[i];
            for (int i = 0; i < nv; i++) qj[i] = qrel[i] * qorg[i];
            eval_f(o, dt, qj, qj_old, a_diag, Rni, rt, &scal_f);
        }
        for (int i = 0; i < nv; i++) qj[i] = qrel[i] * qorg[i];
        mx = 0.0;
        for (int i = 0; i < nv; i++) mx = fmax_f(mx, fabs(rt[i]));
        if (mx < TOLF) break;
        double test = 0.0;
        for (int i = 0; i < nv; i++) {
            double tmp = fabs(qrel[i] - xold[i]) / fmax_f(fabs(qrel[i]), 1.0);
            if (i == 0 || tmp > test) test = tmp;
        }
        if (test < TOLX) break;
    }
    {
        const long n = it > 100 ? 100 : it;
        g_wk.newton_calls++;
        g_wk.newton_its += n;
        if (n > g_wk.newton_max_its) g_wk.newton_max_its = n;
        const int b = n <= 4 ? (int)n - 1 : (n <= 8 ? 4 : (n <= 16 ? 5 : (n <= 32 ? 6 : 7)));
        g_wk.hist[b]++;
    }
    return it;
}

/* imex_RK_solver (N_RK = 2 tableau, src/solver_2d.f90:429-439):
 * src/solver_2d.f90:1022-1634 */
static void imex_RK_solver(oracle_t *o)
{
    const int nv = o->nv, np = nv + 2, ns = o->ns;
    const size_t N = (size_t)o->nx * o->ny;
    const double dt = o->dt;
    double *q_rk0 = o->q_rk, *q_rk1 = o->q_rk + N * nv;
    double *qp_rk0 = o->qp_rk, *qp_rk1 = o->qp_rk + N * np;
    double *dF0 = o->divF, *dF1 = o->divF + N * nv;
    double *NH0 = o->NH, *NH1 = o->NH + N * nv;
    double *SI0 = o->SI, *SI1 = o->SI + N * nv;
    double *ex0 = o->expl, *ex1 = o->expl + N * nv;
    for (int l = 0; l < o->ncells; l++) {
        const size_t c = CELL(o, o->jc[l], o->kc[l]);
        for (int i = 0; i < nv; i++) {
            o->q0[nv * c + i] = o->q[nv * c + i];
            q_rk0[nv * c + i] = 0.0; q_rk1[nv * c + i] = 0.0;
            dF0[nv * c + i] = 0.0; dF1[nv * c + i] = 0.0;
            NH0[nv * c + i] = 0.0; NH1[nv * c + i] = 0.0;
            SI0[nv * c + i] = 0.0; SI1[nv * c + i] = 0.0;
            ex0[nv * c + i] = 0.0; ex1[nv * c + i] = 0.0;
        }
        for (int i = 0; i < np; i++) { qp_rk0[np * c + i] = 0.0; qp_rk1[np * c + i] = 0.0; }
        qp_rk0[np * c + 3] = o->T_amb;
        qp_rk1[np * c + 3] = o->T_amb;
    }
    for (int stage = 0; stage < 2; stage++) {
        const double a_t[2] = {stage == 1 ? 1.0 : 0.0, 0.0};
        const double a_d[2] = {0.0, 0.0};
        const double a_diag = stage == 1 ? 1.0 : 0.0;
        for (int l = 0; l < o->ncells; l++) {
            const int j = o->jc[l], k = o->kc[l];
            const size_t c = CELL(o, j, k);
            o->xi_cell = cell_xi(o, c);
            double q_guess[8], q_si[8], Rj[8], pd;
            double *q0 = o->q0 + nv * c, *qfv = o->q_fv + nv * c, *qpc = o->qp + np * c;
            double *SIs = stage ? SI1 + nv * c : SI0 + nv * c;
            double *NHs = stage ? NH1 + nv * c : NH0 + nv * c;
            if (stage == 0)
                for (int i = 0; i < nv; i++) q_guess[i] = q0[i];
            for (int i = 0; i < nv; i++) {
                double A = 0.0;
                A = A + (dF0[nv * c + i] - ex0[nv * c + i]) * a_t[0];
                A = A + (dF1[nv * c + i] - ex1[nv * c + i]) * a_t[1];
                double B = 0.0;
                B = B + (NH0[nv * c + i] + SI0[nv * c + i]) * a_d[0];
                B = B + (NH1[nv * c + i] + SI1[nv * c + i]) * a_d[1];
                qfv[i] = q0[i] - dt * (A - B);
            }
            qc_to_qp(o, qfv, qpc, &pd);
            if (a_diag != 0.0) {
                if (qfv[0] > 0.0) {
                    eval_nh_semi_impl_terms(o, 1.0, qpc, SIs);
                    for (int i = 0; i < nv; i++) q_si[i] = qfv[i] + dt * a_diag * SIs[i];
                    if ((qfv[1] * qfv[1] + qfv[2] * qfv[2]) == 0.0) {
                        q_si[1] = 0.0;
                        q_si[2] = 0.0;
                    } else if ((q_si[1] * qfv[1] < 0.0) || (q_si[2] * qfv[2] < 0.0)) {
                        q_si[1] = 0.0;
                        q_si[2] = 0.0;
                    } else {
                        double m = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
                        double d = sqrt(qfv[1] * qfv[1] + qfv[2] * qfv[2]);
                        double a1 = m * qfv[1] / d, a2 = m * qfv[2] / d;
                        q_si[1] = a1;
                        q_si[2] = a2;
                    }
                    for (int i = 0; i < nv; i++) SIs[i] = (q_si[i] - qfv[i]) / (dt * a_diag);
                    for (int i = 0; i < nv; i++) q_guess[i] = q_si[i];
                    for (int i = 0; i < nv; i++) {
                        double A = 0.0 + (dF0[nv * c + i] - ex0[nv * c + i]) * a_t[0];
                        double B = 0.0 + (NH0[nv * c + i] + SI0[nv * c + i]) * a_d[0];
                        Rj[i] = (A - B) - a_diag * SIs[i];
                    }
                    g_wk.cell_evals = 0;
                    solve_rk_step(o, dt, q_guess, q0, a_diag, Rj);
                    if (g_wk.cell_evals > g_wk.cell_evals_max) g_wk.cell_evals_max = g_wk.cell_evals;
                    eval_implicit_terms_r(o, q_guess, NHs);
                    if (q_si[1] * q_si[1] + q_si[2] * q_si[2] == 0.0) {
                        q_guess[1] = 0.0;
                        q_guess[2] = 0.0;
                    } else if ((q_guess[1] * q_si[1] <= 0.0) && (q_guess[2] * q_si[2] <= 0.0)) {
                        q_guess[1] = 0.0;
                        q_guess[2] = 0.0;
                    } else {
                        double m = sqrt(q_guess[1] * q_guess[1] + q_guess[2] * q_guess[2]);
                        double d = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
                        double a1 = m * q_si[1] / d, a2 = m * q_si[2] / d;
                        q_guess[1] = a1;
                        q_guess[2] = a2;
                    }
                } else {
                    for (int i = 0; i < nv; i++) {
                        q_guess[i] = qfv[i];
                        q_si[i] = qfv[i];
                        SIs[i] = 0.0;
                        NHs[i] = 0.0;
                    }
                }
                for (int i = 0; i < nv; i++) NHs[i] = (q_guess[i] - q_si[i]) / (dt * a_diag);
            }
            double *qrk = stage ? q_rk1 + nv * c : q_rk0 + nv * c;
            for (int i = 0; i < nv; i++) qrk[i] = q_guess[i];
            if (stage == 0) {
                /* omega_tilde(1) = 1 > 0 */
                double *qprk = qp_rk0 + np * c;
                if (qrk[0] > 0.0) {
                    qc_to_qp(o, qrk, qprk, &pd);
                } else {
                    for (int i = 0; i < np; i++) qprk[i] = 0.0;
                    qprk[3] = o->T_amb;
                }
                eval_expl_terms(o, o->bx[c], o->by[c], 1.0, 0.0, 0.0, qprk, ex0 + nv * c, o->t,
                                0.0, o->arcp[c], o->arcx[c], o->arcy[c], o->dx * o->dy);
            }
        }
        if (stage == 0) {
            /* stale stage-1 primitive states read by neighbours */
            eval_hyperbolic_terms(o, q_rk0, qp_rk0, dF0);
        }
    }
    /* assemble_sol: src/solver_2d.f90:1342-1630 */
    for (int l = 0; l < o->ncells; l++) {
        const size_t c = CELL(o, o->jc[l], o->kc[l]);
        double *qq = o->q + nv * c, pd;
        for (int i = 0; i < nv; i++) qq[i] = q_rk1[nv * c + i];
        for (int i = 0; i < nv; i++)
            if (isnan(qq[i])) {
                o->err |= 1;
                if (o->nan_cell < 0) o->nan_cell = (long)c;
            }
        if (qq[0] < EPS) {
            if (qq[0] > -1.0e-7) {
                for (int i = 0; i < nv; i++) qq[i] = 0.0;
            } else {
                o->err |= 2;
            }
        }
        int anyneg = 0, anybad = 0;
        for (int i = 0; i < ns; i++) {
            if (qq[4 + i] < 0.0) anyneg = 1;
            if (qq[4 + i] <= -1.0e-7) anybad = 1;
        }
        if (anyneg) {
            if (anybad)
                o->err |= 4;
            else
                for (int i = 0; i < ns; i++)
                    if (qq[4 + i] > -1.0e-7) qq[4 + i] = 0.0;
        }
        qc_to_qp(o, qq, o->qp + np * c, &pd);
        if (o->qp[np * c + 3] < 273.0) o->err |= 8;
        double S = 0.0;
        for (int i = 0; i < ns; i++) S += qq[4 + i];
        if (S > qq[0]) {
            double S2 = 0.0;
            for (int i = 0; i < ns; i++) S2 += qq[4 + i];
            int ok = ((S2 - qq[0]) / qq[0] < 1.0e-10) || (qq[0] < EPS);
            if (ok)
                qc_to_qp(o, o->q0 + nv * c, o->qp + np * c, &pd);
            else
                o->err |= 64; /* warning only: the reference rescales and goes on */
            double S3 = 0.0;
            for (int i = 0; i < ns; i++) S3 += qq[4 + i];
            for (int i = 0; i < ns; i++) qq[4 + i] = qq[4 + i] / S3 * qq[0];
        }
    }
    o->xi_cell = o->xi;
}

/* update_erosion_deposition_cell: src/solver_2d.f90:2382-2597 */
static void update_erosion_deposition_cell(oracle_t *o, double dt)
{
    const int nv = o->nv, np = nv + 2, ns = o->ns;
    const size_t N = (size_t)o->nx * o->ny;
    if (o->erosion_coeff == 0.0 && !o->settling && !o->pore && !o->entrainment) return;
    for (int l = 0; l < o->ncells; l++) {
        const size_t c = CELL(o, o->jc[l], o->kc[l]);
        double *qq = o->q + nv * c, *qpc = o->qp + np * c, pd;
        if (qq[0] > 0.0) {
            qc_to_qp(o, qq, qpc, &pd);
        } else {
            for (int i = 0; i < np; i++) qpc[i] = 0.0;
            qpc[3] = o->T_amb;
        }
        double dep[3], eq[8], ers[3], erod[3], topo;
        for (int i = 0; i < ns; i++) erod[i] = o->erodible[c + N * i];
        if (eval_mass_exchange_terms(o, qpc, erod, dt, dep, eq, ers, &topo)) o->err |= 16;
        for (int i = 0; i < nv; i++) qq[i] = qq[i] + dt * eq[i];
        for (int i = 0; i < ns; i++) qq[4 + i] = fmax_f(0.0, qq[4 + i]);
        for (int i = 0; i < ns; i++) o->deposit[c + N * i] = o->deposit[c + N * i] + dt * dep[i];
        for (int i = 0; i < ns; i++) o->erosion[c + N * i] = o->erosion[c + N * i] + dt * ers[i];
        for (int i = 0; i < ns; i++) o->erodible[c + N * i] = o->erodible[c + N * i] - dt * ers[i];
        if (o->erodible_deposit)
            for (int i = 0; i < ns; i++)
                o->erodible[c + N * i] = o->erodible[c + N * i] + dt * dep[i];
        if (o->topo_change) o->bcent[c] = o->bcent[c] + dt * topo; /* src/solver_2d.f90:2493-2497 */
        for (int i = 0; i < ns; i++)
            if (qq[4 + i] < 0.0) o->err |= 4;
        if (qq[0] <= 0.0) {
            if (qq[0] > -1.0e-10) {
                for (int i = 0; i < nv; i++) qq[i] = 0.0;
            } else {
                o->err |= 2;
            }
        }
        double S = 0.0;
        for (int i = 0; i < ns; i++) S += qq[4 + i];
        if (S > qq[0]) {
            if (qq[0] < 1.0e-10) {
                double S2 = 0.0;
                for (int i = 0; i < ns; i++) S2 += qq[4 + i];
                for (int i = 0; i < ns; i++) qq[4 + i] = qq[4 + i] / S2 * qq[0];
            } else {
                o->err |= 32;
            }
        }
        double rg;
        if (qq[0] > 0.0) {
            qc_to_qp(o, qq, qpc, &pd);
            double Ri, rm, rc, c1, c2;
            mixt_var(o, qpc, &Ri, &rm, &rc, &rg, 0, &c1, &c2);
        } else {
            for (int i = 0; i < np; i++) qpc[i] = 0.0;
            qpc[3] = o->T_amb;
            rg = 0.0;
        }
        if (rg <= 0.0)
            for (int i = 0; i < nv; i++) qq[i] = 0.0;
    }
}

/* topography_reconstruction's bed slopes (src/geometry_2d.f90:744-836):
 * the 5-point stencil at the interior cells (1-based 3..n-2), then the
 * copies to the two-cell border, columns first, then rows. */
static void topography_reconstruction(oracle_t *o)
{
    static const double c1[5] = {-2.0, -1.0, 0.0, 1.0, 2.0};
    const int nx = o->nx, ny = o->ny;
    const double n1x = 10.0 * o->dx, n1y = 10.0 * o->dy;
    for (int k = 2; k <= ny - 3; k++)
        for (int j = 2; j <= nx - 3; j++) {
            double sx = 0.0, sy = 0.0;
            for (int m = 0; m < 5; m++) sx = sx + c1[m] * o->bcent[CELL(o, j + m - 2, k)];
            for (int m = 0; m < 5; m++) sy = sy + c1[m] * o->bcent[CELL(o, j, k + m - 2)];
            o->bx[CELL(o, j, k)] = sx / n1x;
            o->by[CELL(o, j, k)] = sy / n1y;
        }
    double *a[2] = {o->bx, o->by};
    for (int m = 0; m < 2; m++) {
        double *b = a[m];
        for (int k = 0; k < ny; k++) {
            b[CELL(o, 0, k)] = b[CELL(o, 2, k)];
            b[CELL(o, 1, k)] = b[CELL(o, 2, k)];
            b[CELL(o, nx - 2, k)] = b[CELL(o, nx - 3, k)];
            b[CELL(o, nx - 1, k)] = b[CELL(o, nx - 3, k)];
        }
        for (int j = 0; j < nx; j++) {
            b[CELL(o, j, 0)] = b[CELL(o, j, 2)];
            b[CELL(o, j, 1)] = b[CELL(o, j, 2)];
            b[CELL(o, j, ny - 2)] = b[CELL(o, j, ny - 3)];
            b[CELL(o, j, ny - 1)] = b[CELL(o, j, ny - 3)];
        }
    }
}

/* driver epilogue: src/IMEX_SfloW2D.f90:449-505 (first=1: :306-353) */
static void epilogue(oracle_t *o, int first)
{
    const int nv = o->nv, np = nv + 2;
    const int nl = o->n_thk * o->n_pdyn;
    for (int l = 0; l < o->ncells; l++) {
        const size_t c = CELL(o, o->jc[l], o->kc[l]);
        double *qpc = o->qp + np * c;
        if (o->q[nv * c] > 0.0) {
            double pd;
            qc_to_qp(o, o->q + nv * c, qpc, &pd);
            if (first)
                o->hmax[c] = qpc[0];
            else
                o->hmax[c] = fmax_f(o->hmax[c], qpc[0]);
            double ru = qpc[nv], rv = qpc[nv + 1];
            double mv2 = ru * ru + rv * rv;
            double mv = sqrt(mv2);
            if (qpc[0] > 0.001) {
                if (first) {
                    o->pdynmax[c] = pd;
                    o->modvelmax[c] = mv;
                } else {
                    o->pdynmax[c] = fmax_f(o->pdynmax[c], pd);
                    o->modvelmax[c] = fmax_f(o->modvelmax[c], mv);
                }
            }
            int it = 0;
            for (int a = 0; a < o->n_thk; a++) {
                int thk = (qpc[0] >= o->thk[a]);
                for (int b = 0; b < o->n_pdyn; b++) {
                    int pdn = (pd >= o->pdyn[b]);
                    if (first)
                        o->vuln[nl * c + it] = (unsigned char)(thk && pdn);
                    else
                        o->vuln[nl * c + it] |= (unsigned char)(thk && pdn);
                    it++;
                }
            }
        } else {
            int n = first ? nv : np;
            for (int i = 0; i < n; i++) qpc[i] = 0.0;
            qpc[3] = o->T_amb;
            if (first) {
                o->hmax[c] = 0.0;
                o->pdynmax[c] = 0.0;
            }
        }
    }
}

/* ====================================================================== */
/* API                                                                     */
/* ====================================================================== */
static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

oracle_t *oracle_create(const int *ip, const double *rp)
{
    oracle_t *o = (oracle_t *)zalloc(sizeof(oracle_t));
    o->nan_cell = -1;
    o->watch_j = o->watch_k = -10;
    o->nx = ip[0]; o->ny = ip[1]; o->ns = ip[2];
    /* pore pressure: ip[16..20], rp[73..78] */
    o->pore = ip[16]; o->gas_loss = ip[17]; o->finh = ip[18]; o->N_inh = ip[19];
    o->dyn_perm = ip[20];
    o->hyd_perm = rp[73]; o->alpha_trans = rp[74];
    o->max_pack = (rp[75] > 0.0) ? rp[75] : 0.64;
    for (int i = 0; i < 3; i++) o->sph_s[i] = (rp[76 + i] > 0.0) ? rp[76 + i] : 1.0;
    o->stoch_tr = ip[54];
    o->nv = 4 + o->ns + (o->stoch_tr ? 1 : 0) + (o->pore ? 1 : 0);
    o->bc_custom = ip[21];
    for (int sd = 0; sd < 4; sd++)
        for (int i = 0; i < 8; i++) {
            o->bcf[sd][i] = o->bc_custom ? ip[22 + 8 * sd + i] : 1;
            o->bcv[sd][i] = o->bc_custom ? rp[79 + 8 * sd + i] : 0.0;
        }
    o->ist = 4 + o->ns;
    o->ipp = 4 + o->ns + (o->stoch_tr ? 1 : 0);
    /* precompute_pascal_coefficient: src/constitutive_2d.f90:419-466 */
    for (int n = 0; n <= o->N_inh && n < 64; n++) {
        const int N = o->N_inh;
        double p1 = 1.0, p2 = 1.0;
        for (int k = 0; k < n; k++) p1 = p1 * (-(double)N - 1.0 - (double)k) / (double)(k + 1);
        for (int k = 0; k < N - n; k++) p2 = p2 * (2.0 * (double)N + 1.0 - (double)k) / (double)(k + 1);
        o->pascal[n] = p1 * p2;
    }
    o->energy = ip[3]; o->settling = ip[4]; o->entrainment = ip[5];
    o->radial = ip[6]; o->collapse = ip[7]; o->n_thk = ip[8]; o->n_pdyn = ip[9];
    o->limiter = ip[10]; o->bcflag = ip[11];
    o->dx = rp[0]; o->dy = rp[0]; o->x0 = rp[1]; o->y0 = rp[2];
    o->dx2 = o->dx / 2.0; o->dy2 = o->dy / 2.0;
    o->one_by_dx = 1.0 / o->dx; o->one_by_dy = 1.0 / o->dy;
    o->cfl = rp[3]; o->dt0 = rp[4]; o->max_dt = rp[5]; o->t_start = rp[6]; o->t_end = rp[7];
    o->grav = rp[8]; o->mu = rp[9]; o->xi = rp[10]; o->pres = rp[11];
    o->inv_pres = 1.0 / o->pres; o->T_amb = rp[12]; o->cp_a = rp[13]; o->R_a = rp[14];
    o->kin_visc_c = rp[15];
    o->rho_a_amb = o->pres / (o->R_a * o->T_amb);
    {
        double d4 = pow(o->dx, 4.0);
        o->eps_sing = fmin_f(fmin_f(d4, d4), 1.0e-10);
        double e = o->eps_sing;
        o->eps_sing4 = e * (e * (e * e));
    }
    o->vel_source = rp[24]; o->T_source = rp[25]; o->h_source = rp[26];
    for (int i = 0; i < 4; i++) o->time_param[i] = rp[27 + i];
    for (int i = 0; i < o->ns; i++) {
        o->rho_s[i] = rp[31 + i]; o->diam_s[i] = rp[34 + i]; o->cp_s[i] = rp[37 + i];
        o->alphas_collapse[i] = rp[40 + i]; o->alphas_source[i] = rp[43 + i];
        o->inv_rho_s[i] = 1.0 / o->rho_s[i];
    }
    for (int i = 0; i < 8; i++) { o->thk[i] = rp[46 + i]; o->pdyn[i] = rp[54 + i]; }
    o->erodible_deposit = ip[12];
    o->topo_change = ip[13];
    o->rheology = (ip[14] == 0) ? 1 : ip[14];
    o->mu_s = rp[69]; o->mu_2 = rp[70]; o->muI_inf = rp[71]; o->I_0 = rp[72];
    o->stochastic = ip[15];
    o->xi_cell = o->xi;
    o->erosion_coeff = rp[63];
    if (o->erosion_coeff == 0.0) { /* src/inpout_2d.f90:1827-1832 */
        o->erodible_porosity = 0.0;
        o->coeff_porosity = 0.0;
        for (int i = 0; i < o->ns; i++) o->erodible_fract[i] = 1.0 / o->ns;
        o->T_erodible = 300.0;
    } else {               /* :1835-1890, the fractions as read_param leaves them */
        o->erodible_porosity = rp[64];
        o->coeff_porosity = o->erodible_porosity / (1.0 - o->erodible_porosity);
        for (int i = 0; i < o->ns; i++) o->erodible_fract[i] = rp[66 + i];
        o->T_erodible = rp[65];
    }
    o->rho_c_sub = o->pres / (o->R_a * o->T_erodible);
    o->pi_g = 4.0 * atan(1.0);
    o->hcs = o->nv * EPS;
    o->one_by_hcs = 1.0 / o->hcs;
    const int nx = o->nx, ny = o->ny, nv = o->nv, np = nv + 2;
    const size_t N = (size_t)nx * ny, NFX = (size_t)(nx + 1) * ny, NFY = (size_t)nx * (ny + 1);
    o->x_stag = zalloc(sizeof(double) * (nx + 1));
    o->y_stag = zalloc(sizeof(double) * (ny + 1));
    o->x_comp = zalloc(sizeof(double) * nx);
    o->y_comp = zalloc(sizeof(double) * ny);
    for (int j = 0; j <= nx; j++) o->x_stag[j] = o->x0 + j * o->dx;
    for (int k = 0; k <= ny; k++) o->y_stag[k] = o->y0 + k * o->dy;
    for (int j = 0; j < nx; j++) o->x_comp[j] = 0.5 * (o->x_stag[j] + o->x_stag[j + 1]);
    for (int k = 0; k < ny; k++) o->y_comp[k] = 0.5 * (o->y_stag[k] + o->y_stag[k + 1]);
    o->bx = zalloc(8 * N); o->by = zalloc(8 * N);
    o->arcp = zalloc(8 * N); o->arcx = zalloc(8 * N); o->arcy = zalloc(8 * N);
    o->scell = zalloc(sizeof(int) * N);
    o->sE = zalloc(N); o->sW = zalloc(N); o->sN = zalloc(N); o->sS = zalloc(N);
    o->q = zalloc(8 * N * nv); o->qp = zalloc(8 * N * np);
    o->q0 = zalloc(8 * N * nv); o->q_fv = zalloc(8 * N * nv);
    o->q_rk = zalloc(8 * N * nv * 2); o->qp_rk = zalloc(8 * N * np * 2);
    o->divF = zalloc(8 * N * nv * 2); o->NH = zalloc(8 * N * nv * 2);
    o->SI = zalloc(8 * N * nv * 2); o->expl = zalloc(8 * N * nv * 2);
    o->qL = zalloc(8 * NFX * nv); o->qR = zalloc(8 * NFX * nv);
    o->qpL = zalloc(8 * NFX * np); o->qpR = zalloc(8 * NFX * np);
    o->qB = zalloc(8 * NFY * nv); o->qT = zalloc(8 * NFY * nv);
    o->qpB = zalloc(8 * NFY * np); o->qpT = zalloc(8 * NFY * np);
    o->axn = zalloc(8 * NFX); o->axp = zalloc(8 * NFX); o->axmax = zalloc(8 * NFX);
    o->ayn = zalloc(8 * NFY); o->ayp = zalloc(8 * NFY); o->aymax = zalloc(8 * NFY);
    o->Hx = zalloc(8 * NFX * nv); o->Hy = zalloc(8 * NFY * nv);
    o->mask = zalloc(N); o->mtmp = zalloc(N); o->mx = zalloc(NFX); o->my = zalloc(NFY);
    o->jc = zalloc(sizeof(int) * N); o->kc = zalloc(sizeof(int) * N);
    o->jsx = zalloc(sizeof(int) * NFX); o->ksx = zalloc(sizeof(int) * NFX);
    o->jsy = zalloc(sizeof(int) * NFY); o->ksy = zalloc(sizeof(int) * NFY);
    o->hmax = zalloc(8 * N); o->pdynmax = zalloc(8 * N); o->modvelmax = zalloc(8 * N);
    o->deposit = zalloc(8 * N * o->ns);
    o->erodible = zalloc(8 * N * o->ns);
    o->bcent = zalloc(8 * N);
    o->Z = zalloc(8 * N);
    o->erosion = zalloc(8 * N * o->ns);
    o->vuln = zalloc(N * (size_t)(o->n_thk * o->n_pdyn > 0 ? o->n_thk * o->n_pdyn : 1));
    o->stL = zalloc(sizeof(int) * NFX); o->stR = zalloc(sizeof(int) * NFX);
    o->stB = zalloc(sizeof(int) * NFY); o->stT = zalloc(sizeof(int) * NFY);
    o->stQP = zalloc(sizeof(int) * N);
    o->stale_cell = zalloc(N);
    /* boundary ring of the mask (src/solver_2d.f90:351-354) */
    for (int j = 0; j < nx; j++) { o->mask[CELL(o, j, 0)] = 1; o->mask[CELL(o, j, ny - 1)] = 1; }
    for (int k = 0; k < ny; k++) { o->mask[CELL(o, 0, k)] = 1; o->mask[CELL(o, nx - 1, k)] = 1; }
    /* allocate_solver_variables: qp = 0, qp(4) = T_ambient (src/solver_2d.f90:286-288) */
    for (size_t c = 0; c < N; c++) o->qp[np * c + 3] = o->T_amb;
    o->t = o->t_start;
    return o;
}

void oracle_destroy(oracle_t *o)
{
    if (!o) return;
    free(o->Z);
    void *ptrs[] = {o->x_comp, o->x_stag, o->y_comp, o->y_stag, o->bx, o->by, o->arcp, o->arcx,
                    o->arcy, o->scell, o->sE, o->sW, o->sN, o->sS, o->q, o->qp, o->q0, o->q_fv,
                    o->q_rk, o->qp_rk, o->divF, o->NH, o->SI, o->expl, o->qL, o->qR, o->qB,
                    o->qT, o->qpL, o->qpR, o->qpB, o->qpT, o->axn, o->axp, o->ayn, o->ayp,
                    o->axmax, o->aymax, o->Hx, o->Hy, o->mask, o->mtmp, o->mx, o->my, o->jc,
                    o->kc, o->jsx, o->ksx, o->jsy, o->ksy, o->hmax, o->pdynmax, o->modvelmax,
                    o->deposit, o->vuln, o->stL, o->stR, o->stB, o->stT, o->stQP, o->stale_cell,
                    o->erodible, o->erosion, o->bcent};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(o);
}

void oracle_set_static(oracle_t *o, const double *bprime_x, const double *bprime_y,
                       const int *source_cell, const int *sides, const double *arc_perim,
                       const double *arc_nx, const double *arc_ny, double h_source)
{
    const size_t N = (size_t)o->nx * o->ny;
    memcpy(o->bx, bprime_x, 8 * N);
    memcpy(o->by, bprime_y, 8 * N);
    for (size_t c = 0; c < N; c++) {
        o->scell[c] = source_cell ? source_cell[c] : 0;
        int s = sides ? sides[c] : 0;
        o->sE[c] = s & 1; o->sW[c] = (s >> 1) & 1; o->sN[c] = (s >> 2) & 1; o->sS[c] = (s >> 3) & 1;
        o->arcp[c] = arc_perim ? arc_perim[c] : 0.0;
        o->arcx[c] = arc_nx ? arc_nx[c] : 0.0;
        o->arcy[c] = arc_ny ? arc_ny[c] : 0.0;
    }
    o->h_source = h_source;
}

void oracle_init_state(oracle_t *o, const double *q)
{
    const size_t N = (size_t)o->nx * o->ny;
    memcpy(o->q, q, 8 * N * o->nv);
    check_solve(o, 1);
    o->dt = o->dt0;
    o->dt_old = o->dt0;
    o->dt_old_old = o->dt0;
    memset(o->hmax, 0, 8 * N);
    memset(o->pdynmax, 0, 8 * N);
    memset(o->modvelmax, 0, 8 * N);
    epilogue(o, 1);
}

int oracle_run(oracle_t *o, int n, double *dts, int *ncells)
{
    const int nv = o->nv;
    const size_t N = (size_t)o->nx * o->ny;
    unsigned char *wet = zalloc(N);
    o->err = 0;
    o->nan_cell = -1;
    for (int s = 0; s < n; s++) {
        for (size_t c = 0; c < N; c++) wet[c] = o->q[nv * c] > 0.0;
        check_solve(o, 0);
        if (ncells) ncells[s] = o->ncells;
        timestep(o);
        if (o->t + o->dt > o->t_end) o->dt = o->t_end - o->t;
        o->dt = fmin_f(o->dt, 1.1 * 0.5 * (o->dt_old + o->dt_old_old));
        o->dt_old_old = o->dt_old;
        o->dt_old = o->dt;
        if (dts) dts[s] = o->dt;
        imex_RK_solver(o);
        update_erosion_deposition_cell(o, o->dt);
        if (o->topo_change) topography_reconstruction(o);
        o->t = o->t + o->dt;
        epilogue(o, 0);
        for (size_t c = 0; c < N; c++)
            if (wet[c] && o->q[nv * c] == 0.0) o->dried++;
    }
    free(wet);
    return o->err;
}

void oracle_get_q(const oracle_t *o, double *q) { memcpy(q, o->q, 8 * (size_t)o->nx * o->ny * o->nv); }
void oracle_set_q(oracle_t *o, const double *q) { memcpy(o->q, q, 8 * (size_t)o->nx * o->ny * o->nv); }

void oracle_get_fields(const oracle_t *o, double *hmax, double *pdynmax, double *modvelmax,
                       double *deposit, double *t, double *dt)
{
    const size_t N = (size_t)o->nx * o->ny;
    memcpy(hmax, o->hmax, 8 * N);
    memcpy(pdynmax, o->pdynmax, 8 * N);
    memcpy(modvelmax, o->modvelmax, 8 * N);
    memcpy(deposit, o->deposit, 8 * N * o->ns);
    *t = o->t;
    *dt = o->dt;
}

void oracle_get_vuln(const oracle_t *o, int *v)
{
    const size_t N = (size_t)o->nx * o->ny;
    const int nl = o->n_thk * o->n_pdyn > 0 ? o->n_thk * o->n_pdyn : 1;
    for (size_t i = 0; i < N * nl; i++) v[i] = (o->n_thk * o->n_pdyn > 0) ? o->vuln[i] : 0;
}

void oracle_get_mask(const oracle_t *o, int *m)
{
    for (size_t c = 0; c < (size_t)o->nx * o->ny; c++) m[c] = o->mask[c];
}

void oracle_set_pow_mode(oracle_t *o, int mode) { o->pow_mode = mode; }

/* 1: read the states that the reference keeps from an earlier step (for dry
 * cells that left the active set) as a fresh evaluation gives them, as the
 * device does; 0 (default): as the reference reads them */
void oracle_set_fresh(oracle_t *o, int fresh) { o->fresh = fresh; }

void oracle_set_time(oracle_t *o, double t, double dt_old, double dt_old_old)
{
    o->t = t;
    o->dt_old = dt_old;
    o->dt_old_old = dt_old_old;
}

/* The end time of the driver's dt clip (an output time, src/IMEX_SfloW2D.f90:417-433). */
void oracle_set_t_end(oracle_t *o, double t_end) { o->t_end = t_end; }

double oracle_get_t(const oracle_t *o) { return o->t; }

long oracle_get_nan_cell(const oracle_t *o) { return o->nan_cell; }

void oracle_watch_cell(oracle_t *o, int j, int k)
{
    o->watch_j = j;
    o->watch_k = k;
    o->watch_seen[0] = o->watch_seen[1] = 0;
}

int oracle_watched_faces(const oracle_t *o, double *pre, double *post)
{
    for (int f = 0; f < 2; f++)
        for (int i = 0; i < 8; i++) {
            pre[8 * f + i] = o->watch_pre[f][i];
            post[8 * f + i] = o->watch_post[f][i];
        }
    return o->watch_seen[0] + 2 * o->watch_seen[1];
}

void oracle_work_reset(void) { memset(&g_wk, 0, sizeof g_wk); }

void oracle_work_stats(long *out)
{
    const long v[18] = {g_wk.newton_calls, g_wk.newton_its, g_wk.newton_max_its, g_wk.evalf,
                        g_wk.jac, g_wk.ls_trials, g_wk.cell_evals_max, g_wk.settle_calls,
                        g_wk.settle_its, g_wk.settle_max_its, g_wk.hist[0], g_wk.hist[1],
                        g_wk.hist[2], g_wk.hist[3], g_wk.hist[4], g_wk.hist[5], g_wk.hist[6],
                        g_wk.hist[7]};
    memcpy(out, v, sizeof v);
}

/* 1 for every cell next to a stale read that differs from a fresh
 * evaluation since the handle was created (the cells whose results the
 * reference's stale-state behaviour can change) */
/* the stage arrays of cell (j, k) after the last step, as the harness's
 * imexref_get_stage lays them out: [5][n_RK][nv] = q_rk, divF, NH, SI, expl */
void oracle_get_stage(const oracle_t *o, int j, int k, double *out)
{
    const size_t N = (size_t)o->nx * o->ny, c = CELL(o, j, k);
    const int nv = o->nv;
    const double *src[5] = {o->q_rk, o->divF, o->NH, o->SI, o->expl};
    for (int a = 0; a < 5; a++)
        for (int st = 0; st < 2; st++)
            for (int i = 0; i < nv; i++) out[(a * 2 + st) * nv + i] = src[a][N * nv * st + nv * c + i];
}

void oracle_get_stale_cells(const oracle_t *o, unsigned char *out)
{
    memcpy(out, o->stale_cell, (size_t)o->nx * o->ny);
}

void oracle_get_counters(const oracle_t *o, long *dried, long *stale)
{
    *dried = o->dried;
    *stale = o->stale;
}

void oracle_qc_to_qp(const oracle_t *o, int n, const double *qc, double *qp, double *pdyn)
{
    for (int i = 0; i < n; i++) qc_to_qp(o, qc + (size_t)o->nv * i, qp + (size_t)(o->nv + 2) * i, pdyn + i);
}
void oracle_qp_to_qc(const oracle_t *o, int n, const double *qp, double *qc)
{
    for (int i = 0; i < n; i++) qp_to_qc(o, qp + (size_t)(o->nv + 2) * i, qc + (size_t)o->nv * i);
}
void oracle_mixt_var(const oracle_t *o, int n, const double *qp, double *out)
{
    for (int i = 0; i < n; i++) {
        double *r = out + 6 * (size_t)i;
        mixt_var(o, qp + (size_t)(o->nv + 2) * i, r, r + 1, r + 2, r + 3, 1, r + 4, r + 5);
    }
}
void oracle_fluxes(const oracle_t *o, int n, int dir, const double *qc, const double *qp, double *fl)
{
    for (int i = 0; i < n; i++)
        eval_fluxes(o, qc + (size_t)o->nv * i, qp + (size_t)(o->nv + 2) * i, 1.0, dir, fl + (size_t)o->nv * i);
}
void oracle_speeds(const oracle_t *o, int n, int dir, const double *qp, double *out)
{
    for (int i = 0; i < n; i++)
        eval_local_speeds(o, qp + (size_t)(o->nv + 2) * i, 1.0, dir, out + 2 * i, out + 2 * i + 1);
}
void oracle_implicit(const oracle_t *o, int n, const double *qc, double *nh)
{
    for (int i = 0; i < n; i++) eval_implicit_terms_r(o, qc + (size_t)o->nv * i, nh + (size_t)o->nv * i);
}
void oracle_implicit_cplx(const oracle_t *o, int n, const double *qre, const double *qim,
                          double *ore, double *oim)
{
    const int nv = o->nv;
    for (int i = 0; i < n; i++) {
        cx qj[8], r[8];
        for (int m = 0; m < nv; m++) qj[m] = C(qre[nv * i + m], qim[nv * i + m]);
        eval_implicit_terms_c(o, qj, r);
        for (int m = 0; m < nv; m++) { ore[nv * i + m] = r[m].re; oim[nv * i + m] = r[m].im; }
    }
}
void oracle_semi_impl(const oracle_t *o, int n, const double *qc, const double *qp, double *si)
{
    (void)qc;
    for (int i = 0; i < n; i++) eval_nh_semi_impl_terms(o, 1.0, qp + (size_t)(o->nv + 2) * i, si + (size_t)o->nv * i);
}
void oracle_expl(const oracle_t *o, int n, const double *bx, const double *by, double t,
                 const double *qp, double *ex)
{
    for (int i = 0; i < n; i++)
        eval_expl_terms(o, bx[i], by[i], 1.0, 0.0, 0.0, qp + (size_t)(o->nv + 2) * i, ex + (size_t)o->nv * i, t,
                        0.0, 0.0, 0.0, 0.0, o->dx * o->dy);
}
void oracle_settling(const oracle_t *o, int n, const double *diam, const double *rhos,
                     const double *rhoc, const double *ikv, double *out)
{
    for (int i = 0; i < n; i++) out[i] = settling_velocity(o, diam[i], rhos[i], rhoc[i], ikv[i]);
}
void oracle_mass_exchange(const oracle_t *o, int n, const double *qp, double dt, double *dep,
                          double *eq)
{
    const double zero[3] = {0.0, 0.0, 0.0};
    double ers[3];
    for (int i = 0; i < n; i++) {
        double topo;
        eval_mass_exchange_terms(o, qp + (size_t)(o->nv + 2) * i, zero, dt, dep + (size_t)o->ns * i,
                                 eq + (size_t)o->nv * i, ers, &topo);
    }
}
/* with the cells' erodible layers erod[n][n_solid]; ers[n][n_solid] the erosion rates */
void oracle_mass_exchange_ers(const oracle_t *o, int n, const double *qp, const double *erod,
                              double dt, double *dep, double *eq, double *ers)
{
    for (int i = 0; i < n; i++) {
        double topo;
        eval_mass_exchange_terms(o, qp + (size_t)(o->nv + 2) * i, erod + (size_t)o->ns * i, dt,
                                 dep + (size_t)o->ns * i, eq + (size_t)o->nv * i,
                                 ers + (size_t)o->ns * i, &topo);
    }
}
/* B_cent (cells x fastest), evolved with topo_change_flag */
void oracle_set_bcent(oracle_t *o, const double *b)
{
    memcpy(o->bcent, b, 8 * (size_t)o->nx * o->ny);
}
void oracle_get_bcent(const oracle_t *o, double *b)
{
    memcpy(b, o->bcent, 8 * (size_t)o->nx * o->ny);
}
/* erodible[class][cell] (cells x fastest) */
void oracle_set_erodible(oracle_t *o, const double *e)
{
    memcpy(o->erodible, e, 8 * (size_t)o->nx * o->ny * o->ns);
}
/* Z(nx, ny), cells x fastest */
void oracle_set_noise(oracle_t *o, const double *Z)
{
    memcpy(o->Z, Z, 8 * (size_t)o->nx * o->ny);
}
void oracle_get_erosion(const oracle_t *o, double *ers, double *erd)
{
    memcpy(ers, o->erosion, 8 * (size_t)o->nx * o->ny * o->ns);
    memcpy(erd, o->erodible, 8 * (size_t)o->nx * o->ny * o->ns);
}
double oracle_limit_minmod(const double *v3, const double *z3) { return limit_minmod(v3, z3); }

int oracle_newton(const oracle_t *o, double dt, const double *q_si, const double *q0,
                  const double *R, double *q)
{
    for (int i = 0; i < o->nv; i++) q[i] = q_si[i];
    return solve_rk_step(o, dt, q, q0, 1.0, R);
}
