// Per-cell and per-face physics of the IMEX-RK stage, as device functions.
//
// Every function evaluates the same floating-point operations in the same
// order as the reference routine it is named after (cited per function), so
// the FP64 instantiation reproduces the reference bit for bit wherever only
// +, -, *, / and sqrt are involved.  This file is compiled with
// -ffp-contract=off: no multiply-add may be fused.
//
// NS = number of solid classes, NV = 4 + NS conservative variables, primitive
// vectors have NV + 2 entries (h, hu, hv, T, h*alpha_i, u, v).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "imex_math.h"

namespace imex {

constexpr double kEps = 2.220446049250313e-16;  // EPSILON(1.0_wp)
constexpr double kSqrt2 = 1.4142135623730951;     // SQRT(2.0_wp)

// Run constants, copied to the device once per handle.
struct Params {
  int nx, ny, ns, nv;
  int energy, settling, entrainment, radial;
  int n_thk, n_pdyn;
  // 1 when a cell of the uploaded state is dry (q1 <= 0) but not all zero:
  // then every tile with active cells goes through k_flux, not only the
  // tiles near wet or source-ring cells
  int all_tiles;
  double dx, dy, dx2, dy2, one_by_dx, one_by_dy;
  double cfl, max_dt;
  double grav, mu, xi, pres, inv_pres, T_amb, cp_a, R_a, kin_visc_c, rho_a_amb;
  double eps_sing, eps_sing4;
  double rho_s[3], inv_rho_s[3], diam_s[3], cp_s[3], alphas_source[3];
  double rho_c_sub, T_erodible;
  // erosion of the erodible layer (src/inpout_2d.f90:1813-1900)
  double erosion_coeff, erodible_porosity, coeff_porosity, erodible_fract[3];
  int erodible_deposit;
  int topo_change;  // B_cent follows erosion and deposition (topo_change_flag)
  // friction law: 1 Voellmy-Salm, 11 mu(I), 12 mu(I) + Voellmy turbulent term
  int rheology;
  // stochastic friction: a per-cell value Z shifts xi (stochastic_flag)
  int stochastic;
  double mu_s, mu_2, muI_inf, I_0;
  double vel_source, T_source, h_source, time_param[4], pi_g;
  double thk[8], pdyn[8];
  double hcs, one_by_hcs;  // complex step n_vars*eps and its reciprocal
  // pore pressure (PORE_PRESSURE_PARAMETERS): the conserved variable
  // q1 * p_exc after the solid classes (and after the transported noise)
  int pore, gas_loss, finh, N_inh, dyn_perm;  // finh: 0 OFF, 1 STEP, 2 STATIC, 3 DYNAMIC
  // the transported noise (stoch_transport_flag): q1 * Z after the solid
  // classes; general: pore or stoch_tr, the runs of imex_pore.hip's path
  int stoch_tr, general;
  double hyd_perm, alpha_trans, max_pack, sph_s[3], pascal[16];
  // per-variable boundary conditions of the pore-pressure path: side 0 W,
  // 1 E, 2 S, 3 N; flag 0 Dirichlet value, 1 slope value, 2 one-sided slope
  int bcf[4][8];
  double bcv[4][8];
};

// Per-cell phase clocks of the update kernel, compiled in only by the
// phase-clock build (-DIMEX_PHASE_CLOCKS; imex_sflow2d_dev_amd/build.py
// --phase-clocks, read by imex_sflow2d_dev_amd/phase_clocks.py): shader-clock
// readings at the phase boundaries of one cell's update and its work counts.
// In every other build the calls below compile to nothing.
#ifdef IMEX_PHASE_CLOCKS
constexpr bool kPhaseClocks = true;
#else
constexpr bool kPhaseClocks = false;
#endif
enum PhaseSlot {
  PC_START = 0,      // the cell's update begins (work-list entry read)
  PC_STAGE2 = 1,     // stage 2 begins (state and right-hand side loaded)
  PC_NEWTON0 = 2,    // the Newton solve begins (q_fv, semi-implicit terms done)
  PC_NEWTON1 = 3,    // the Newton solve ends
  PC_STAGE2END = 4,  // assembly checks done
  PC_SED0 = 5,       // sedimentation begins
  PC_SED1 = 6,       // sedimentation ends
  PC_END = 7,        // the cell's stores are issued
  PC_ITS = 8,        // Newton iterations
  PC_JAC = 9,        // cycles in the complex-step Jacobian
  PC_LNSRCH = 10,    // cycles in the line search
  PC_TRIALS = 11,    // line-search rounds (two trials each on a lane pair)
  PC_SETTLE = 12,    // cycles in the settling-velocity fixed points
  PC_SETTLE_ITS = 13,  // settling fixed-point iterations (this lane's classes)
  PC_POW = 14,       // cycles in the hindered-settling and entrainment powers
  PC_CPLX = 15,      // residual evaluations that took the complex path
  PC_NSLOT = 16
};
struct PhaseClk {
  unsigned long long t[PC_NSLOT];
};
__device__ __forceinline__ unsigned long long pc_now() {
  if constexpr (kPhaseClocks) return clock64();
  return 0ULL;
}
__device__ __forceinline__ void pc_mark(PhaseClk* pc, int slot) {
  if constexpr (kPhaseClocks) {
    if (pc) pc->t[slot] = clock64();
  }
}
__device__ __forceinline__ void pc_add(PhaseClk* pc, int slot, unsigned long long v) {
  if constexpr (kPhaseClocks) {
    if (pc) pc->t[slot] += v;
  }
}

// Fortran MIN(a, b) / MAX(a, b) as the reference build evaluates them:
// (a < b) ? a : b and (a > b) ? a : b, so a NaN or a signed-zero tie
// returns the second argument.
template <typename T>
__device__ __forceinline__ T fmin_f(T a, T b) { return (a < b) ? a : b; }
template <typename T>
__device__ __forceinline__ T fmax_f(T a, T b) { return (a > b) ? a : b; }

// ---------------------------------------------------------------------------
// complex arithmetic as the reference build evaluates it
// ---------------------------------------------------------------------------
struct cx {
  double re, im;
};
__device__ __forceinline__ cx C(double re, double im) { return cx{re, im}; }
__device__ __forceinline__ cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cx cmul(cx a, cx b) {
  return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cx cconj(cx a) { return C(a.re, -a.im); }
__device__ __forceinline__ cx csq(cx a) {
  return C(a.re * a.re - a.im * a.im, a.re * a.im + a.re * a.im);
}

// C99 Annex G.5.1 scaled division (logb / scalbn are exact operations).
__device__ inline cx cdiv(cx z, cx w) {
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

// sqrt(x^2 + y^2) for finite x, y: scaled by a power of two, the sum of
// squares is formed exactly as a double-double (each square split with an
// exact product, Dekker/Knuth two-sum), then rounded once through sqrt and a
// single correction step.  Exact (equal to |x|) whenever |y| is below half an
// ulp of |x| in relative terms.
__device__ inline double hypot_dd(double x, double y) {
  x = fabs(x);
  y = fabs(y);
  if (x < y) {
    double t = x;
    x = y;
    y = t;
  }
  if (x == 0.0 || isinf(x)) return x;
  if (y == 0.0) return x;
  int e = ilogb(x);
  double xs = scalbn(x, -e), ys = scalbn(y, -e);
  // exact squares as (hi, lo) pairs
  double xh = xs * xs, xl = fma(xs, xs, -xh);
  double yh = ys * ys, yl = fma(ys, ys, -yh);
  // two-sum of the high parts
  double s = xh + yh;
  double bb = s - xh;
  double err = (xh - (s - bb)) + (yh - bb);
  double lo = err + xl + yl;
  double r = sqrt(s + lo);
  // one Newton correction: r += (s + lo - r*r) / (2r), r*r formed exactly
  double rr = r * r, rrl = fma(r, r, -rr);
  double resid = ((s - rr) + lo) - rrl;
  r = r + resid / (2.0 * r);
  return scalbn(r, e);
}

// csqrt with the case split of the C library implementation used by the
// reference build: imaginary zero, real zero, and the general formula
// r = sqrt((|z| + |x|)/2), s = y / (2 r).
__device__ inline cx csqrt_c(cx z) {
  double x = z.re, y = z.im;
  if (y == 0.0) {
    if (x < 0.0) return C(0.0, copysign(sqrt(-x), y));
    return C(fabs(sqrt(x)), copysign(0.0, y));
  }
  if (x == 0.0) {
    double r;
    if (fabs(y) >= 2.0 * 2.2250738585072014e-308)
      r = sqrt(0.5 * fabs(y));
    else
      r = 0.5 * sqrt(2.0 * fabs(y));
    return C(r, copysign(r, y));
  }
  double d = hypot_dd(x, y);
  double r, s;
  if (x > 0.0) {
    r = sqrt(0.5 * (d + x));
    s = 0.5 * (y / r);
  } else {
    s = sqrt(0.5 * (d - x));
    r = fabs(0.5 * (y / s));
  }
  return C(r, copysign(s, y));
}

// Fortran COMPLEX**REAL in the reference build: the C library's cpow(z, c +
// 0i) = cexp(c * clog(z)), evaluated here with the same steps and correctly
// rounded functions: clog's real part by the C library's case split (|x| = 1;
// 1 < |x| < 2 with |y| < 1; 0.5 <= |x| < 1 with |y| < eps/2, or with
// x**2 + y**2 >= 0.5; otherwise log(hypot)), its imaginary part atan2(y, x),
// the product with c + 0i as (cL - 0 th, c th + 0 L), and cexp(a + ib) as
// exp(a) (cos b, sin b) with (cos b, sin b) = (1, b) for |b| <= DBL_MIN.
// atan2, sin and cos are correctly rounded for the arguments the solver
// gives them (|y| <= 2**-30 x, x > 0; |b| < 2**-27: y/x, b and 1); others take
// the device library's functions.
__device__ inline cx cpow_c(cx z, double c) {
  double ax = fabs(z.re), ay = fabs(z.im);
  if (ax < ay) {
    const double t = ax;
    ax = ay;
    ay = t;
  }
  double L;
  if (ax == 1.0) {
    L = ddm::log1p_cr(ay * ay) / 2.0;
  } else if (ax > 1.0 && ax < 2.0 && ay < 1.0) {
    double d2m1 = (ax - 1.0) * (ax + 1.0);
    if (ay >= kEps) d2m1 += ay * ay;
    L = ddm::log1p_cr(d2m1) / 2.0;
  } else if (ax < 1.0 && ax >= 0.5 && ay < kEps / 2.0) {
    const double d2m1 = (ax - 1.0) * (ax + 1.0);
    L = ddm::log1p_cr(d2m1) / 2.0;
  } else if (ax < 1.0 && ax >= 0.5 && ax * ax + ay * ay >= 0.5) {
    L = ddm::log1p_cr(ddm::x2y2m1_cr(ax, ay)) / 2.0;
  } else {
    L = ddm::log_cr(hypot_dd(ax, ay)) - 0.0 * 0x1.62e42fefa39efp-1;
  }
  double th;
  if (z.re > 0.0 && fabs(z.im) <= 0x1.0p-30 * z.re)
    th = z.im / z.re;
  else
    th = atan2(z.im, z.re);
  const double a = c * L - 0.0 * th, b = c * th + 0.0 * L;
  double sb, cb;
  if (fabs(b) > 2.2250738585072014e-308) {
    if (fabs(b) < 0x1.0p-27) {
      sb = b;
      cb = 1.0;
    } else {
      sb = sin(b);
      cb = cos(b);
    }
  } else {
    sb = b;
    cb = 1.0;
  }
  const double E = ddm::exp_cr(a);
  return C(E * cb, E * sb);
}

// ===========================================================================
// constitutive_2d
// ===========================================================================

// The functions of the explicit part (primitive variables, local speeds,
// physical fluxes, explicit terms, reconstruction) take the arithmetic type T
// as a template parameter: double everywhere on the FP64 path (then every
// T(...) below is an identity and the operations are the reference's), float
// in the FP32 form's flux kernel.

// NX: the extra conserved variables after the solid classes, as bits
// (0 everywhere but the general path of csrc/imex_pore.hip): 2 the
// transported noise q1 * Z at index 4 + NS (idx_stoch), 1 the pore pressure
// q1 * p_exc after it (idx_pore; src/inpout_2d.f90:642-655)
constexpr int nx_count(int NX) { return (NX & 1) + ((NX >> 1) & 1); }
constexpr bool has_pore(int NX) { return (NX & 1) != 0; }
constexpr bool has_stoch(int NX) { return (NX & 2) != 0; }
template <int NS, int NX>
constexpr int pore_at() { return 4 + NS + ((NX >> 1) & 1); }
template <int NS>
constexpr int stoch_at() { return 4 + NS; }

// r_phys_var (configured branch) + qc_to_qp:
// src/constitutive_2d.f90:550-976 and :1637-1690.  Also returns p_dyn.
template <int NS, typename T = double, int NX = 0>
__device__ inline void qc_to_qp(const Params& P, const T* qc, T* qp, T& p_dyn) {
  constexpr int NV = 4 + NS + nx_count(NX), PI = pore_at<NS, NX>(), SI = stoch_at<NS>();
  T xs[NS];
  T inv_qj1, pexc = T(0.0), zs = T(0.0);
  if (qc[0] > T(kEps)) {
    inv_qj1 = T(1.0) / qc[0];
    if constexpr (has_stoch(NX)) zs = qc[SI] * inv_qj1;  // :657
    if constexpr (has_pore(NX)) pexc = qc[PI] * inv_qj1;  // :659
#pragma unroll
    for (int i = 0; i < NS; i++) xs[i] = qc[4 + i] * inv_qj1;
    T s = T(0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) s += qc[4 + i];
    if (s == qc[0]) {
      T sx = T(0.0);
#pragma unroll
      for (int i = 0; i < NS; i++) sx += xs[i];
#pragma unroll
      for (int i = 0; i < NS; i++) xs[i] = xs[i] / sx;
    }
  } else {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = T(0.0);
    qp[3] = T(P.T_amb);
    p_dyn = T(0.0);
    return;
  }
  T xs_tot = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) xs_tot += xs[i];
  T xc = T(1.0) - xs_tot;
  T cpc_by_xc = (xc - T(0.0)) * T(P.cp_a) + T(0.0);
  T dot = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) dot += xs[i] * T(P.cp_s[i]);
  T cp_mix = dot + cpc_by_xc;
  T Tm;
  if (qc[0] > T(P.eps_sing)) {
    if (P.energy)
      Tm = (qc[3] - T(0.5) * (qc[1] * qc[1] + qc[2] * qc[2]) * inv_qj1) / (qc[0] * cp_mix);
    else
      Tm = qc[3] / (qc[0] * cp_mix);
    if (Tm <= T(0.0)) Tm = T(P.T_amb);
  } else {
    Tm = T(P.T_amb);
  }
  T Rc;
  if (xc > T(kEps))
    Rc = ((xc - T(0.0)) * T(P.R_a) + T(0.0)) / xc;
  else
    Rc = T(P.R_a);
  T inv_rho_c = Rc * Tm * T(P.inv_pres);
  T inv_rhom = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) inv_rhom += xs[i] * T(P.inv_rho_s[i]);
  inv_rhom = inv_rhom + xc * inv_rho_c;
  T rm = T(1.0) / inv_rhom;
  T h = qc[0] * inv_rhom;
  T u, v;
  if (qc[0] > T(P.eps_sing)) {
    u = qc[1] * inv_qj1;
    v = qc[2] * inv_qj1;
  } else {
    T q1 = qc[0];
    T den = sqrt(q1 * (q1 * (q1 * q1)) + T(P.eps_sing4));
    u = T(kSqrt2) * q1 * qc[1] / den;
    v = T(kSqrt2) * q1 * qc[2] / den;
  }
  p_dyn = T(0.5) * rm * (u * u + v * v);
  qp[0] = h;
  qp[1] = h * u;
  qp[2] = h * v;
  qp[3] = Tm;
#pragma unroll
  for (int i = 0; i < NS; i++) qp[4 + i] = xs[i] * rm * T(P.inv_rho_s[i]) * h;
  if constexpr (has_stoch(NX)) qp[SI] = zs;  // :1681
  if constexpr (has_pore(NX)) qp[PI] = pexc;  // :1683
  qp[NV] = u;
  qp[NV + 1] = v;
}

// qp_to_qc (configured branch): src/constitutive_2d.f90:1718-2149
template <int NS, typename T = double, int NX = 0>
__device__ inline void qp_to_qc(const Params& P, const T* qp, T* qc) {
  constexpr int NV = 4 + NS + nx_count(NX), PI = pore_at<NS, NX>(), SI = stoch_at<NS>();
  T h = qp[0];
  if (!(h > T(kEps))) {
#pragma unroll
    for (int i = 0; i < NV; i++) qc[i] = T(0.0);
    return;
  }
  T hu = qp[1], hv = qp[2], u = qp[NV], v = qp[NV + 1], Tm = qp[3];
  T alphas[NS], ar[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) alphas[i] = qp[4 + i] / qp[0];
  T at = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) at += alphas[i];
#pragma unroll
  for (int i = 0; i < NS; i++) ar[i] = alphas[i] * T(P.rho_s[i]);
  T rho_a = T(P.pres) / (T(P.R_a) * Tm);
  T rho_c = (((T(1.0) - T(0.0)) - at - T(0.0)) * rho_a + T(0.0)) / ((T(1.0) - T(0.0)) - at);
  T alphac = T(1.0) - at;
  T sar = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) sar += ar[i];
  T rho_m = alphac * rho_c + sar;
  T inv_rhom = T(1.0) / rho_m;
  T xs[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) xs[i] = ar[i] * inv_rhom;
  T xc = alphac * rho_c * inv_rhom;
  T cp_c;
  if (xc > T(kEps))
    cp_c = ((xc - T(0.0)) * T(P.cp_a) + T(0.0)) / xc;
  else
    cp_c = T(P.cp_a);
  T dot = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) dot += xs[i] * T(P.cp_s[i]);
  T cp_mix = dot + xc * cp_c;
  qc[0] = rho_m * h;
  qc[1] = rho_m * hu;
  qc[2] = rho_m * hv;
  if (P.energy) {
    if (h > T(0.0))
      qc[3] = h * rho_m * (cp_mix * Tm + T(0.5) * (u * u + v * v));
    else
      qc[3] = T(0.0);
  } else {
    qc[3] = h * rho_m * cp_mix * Tm;
  }
#pragma unroll
  for (int i = 0; i < NS; i++) qc[4 + i] = xs[i] * qc[0];
  if constexpr (has_stoch(NX)) qc[SI] = qp[SI] * qc[0];  // :1979, :2141
  if constexpr (has_pore(NX)) qc[PI] = qp[PI] * qc[0];  // :1981, :2143
}

template <typename T = double>
struct MixtT {
  T Ri, rho_m, rho_c, red_grav, cp_c, cp_mix;
};
using Mixt = MixtT<double>;

// mixt_var (+ eval_sp_heat when sp_heat): src/constitutive_2d.f90:1342-1607
template <int NS, typename T = double, int NX = 0>
__device__ inline MixtT<T> mixt_var(const Params& P, const T* qp, bool sp_heat) {
  constexpr int NV = 4 + NS + nx_count(NX);
  MixtT<T> m;
  T h = qp[0];
  if (qp[0] <= T(kEps)) {
    m.red_grav = T(0.0);
    m.rho_m = T(P.rho_a_amb);
    m.rho_c = T(P.rho_a_amb);
    m.cp_c = T(P.cp_a);
    m.cp_mix = T(P.cp_a);
    m.Ri = T(0.0);
    return m;
  }
  T u = qp[NV], v = qp[NV + 1], Tm = qp[3];
  T alphas[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) alphas[i] = qp[4 + i] / qp[0];
  T at = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) at += alphas[i];
  T rho_a = T(P.pres) / (T(P.R_a) * Tm);
  T rc = (((T(1.0) - T(0.0)) - at - T(0.0)) * rho_a + T(0.0)) / ((T(1.0) - T(0.0)) - at);
  T dot = T(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) dot += alphas[i] * T(P.rho_s[i]);
  T rm = (T(1.0) - at) * rc + dot;
  T inv_rhom = T(1.0) / rm;
  T rg = (rm - T(P.rho_a_amb)) / rm * T(P.grav);
  T u2 = u * u + v * v;
  m.Ri = (u2 > T(0.0)) ? fmin_f(T(1.0e15), rg * h / u2) : T(0.0);
  m.rho_m = rm;
  m.rho_c = rc;
  m.red_grav = rg;
  m.cp_c = T(0.0);
  m.cp_mix = T(0.0);
  if (sp_heat) {
    T xs[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) xs[i] = alphas[i] * T(P.rho_s[i]) * inv_rhom;
    T sxs = T(0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) sxs += xs[i];
    T xc = T(1.0) - sxs;
    T c = (xc > T(kEps)) ? ((xc - T(0.0)) * T(P.cp_a) + T(0.0)) / xc : T(P.cp_a);
    T d2 = T(0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) d2 += xs[i] * T(P.cp_s[i]);
    m.cp_c = c;
    m.cp_mix = d2 + xc * c;
  }
  return m;
}

// eval_local_speeds_x / _y with grav_coeff = 1: src/constitutive_2d.f90:2203-2307
template <int NS, typename T = double, int NX = 0>
__device__ inline void local_speeds(const Params& P, const T* qp, int dir, T& vmin, T& vmax) {
  constexpr int NV = 4 + NS + nx_count(NX);
  MixtT<T> m = mixt_var<NS, T, NX>(P, qp, false);
  T h = qp[0];
  T w = (dir == 1) ? qp[NV] : qp[NV + 1];
  T rg = m.red_grav;
  if (rg * h < T(0.0)) {
    vmin = w;
    vmax = w;
  } else {
    T cel = (dir == 1) ? sqrt(rg * h * T(1.0)) : sqrt(T(1.0) * rg * h);
    vmin = w - cel;
    vmax = w + cel;
  }
}

// eval_fluxes with grav_coeff = 1: src/constitutive_2d.f90:2325-2518
template <int NS, typename T = double, int NX = 0>
__device__ inline void phys_flux(const Params& P, const T* qc, const T* qp, int dir, T* f) {
  constexpr int NV = 4 + NS + nx_count(NX);
  if (qp[0] > T(kEps)) {
    T h = qp[0];
    T w = (dir == 1) ? qp[NV] : qp[NV + 1];
    MixtT<T> m = mixt_var<NS, T, NX>(P, qp, false);
    T Pr = T(0.5) * m.rho_m * T(1.0) * m.red_grav * (h * h);
    f[0] = w * qc[0] * T(1.0);
    if (dir == 1) {
      f[1] = w * qc[1] * T(1.0) + Pr;
      f[2] = w * qc[2] * T(1.0);
    } else {
      f[1] = w * qc[1] * T(1.0);
      f[2] = w * qc[2] * T(1.0) + Pr;
    }
    f[3] = P.energy ? w * (qc[3] * T(1.0) + Pr) : w * qc[3] * T(1.0);
#pragma unroll
    for (int i = 0; i < NS; i++) f[4 + i] = w * qc[4 + i] * T(1.0);
    T S = T(0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) S += f[4 + i];
    if ((f[0] > T(0.0)) && (S / f[0] > T(1.0))) {
#pragma unroll
      for (int i = 0; i < NS; i++) f[4 + i] = f[4 + i] / S * f[0];
    }
    // fluxes of q1 * Z and q1 * p_exc: :2428-2439, :2490-2501
    if constexpr (has_stoch(NX)) f[stoch_at<NS>()] = w * qc[stoch_at<NS>()] * T(1.0);
    if constexpr (has_pore(NX)) f[pore_at<NS, NX>()] = w * qc[pore_at<NS, NX>()] * T(1.0);
  } else {
#pragma unroll
    for (int i = 0; i < NV; i++) f[i] = T(0.0);
  }
}

// eval_expl_terms (gravity, lateral radial volume source), grav_coeff = 1,
// no curvature, d_grav_coeff = 0: src/constitutive_2d.f90:3024-3431.  The
// source's time factor and inflow state are computed in double and rounded
// to T.
template <int NS, typename T = double, int NX = 0>
__device__ inline void expl_terms(const Params& P, T Bx, T By, const T* qp, T* ex, double time,
                                  T arc_p, T arc_x, T arc_y) {
  constexpr int NV = 4 + NS + nx_count(NX);
#pragma unroll
  for (int i = 0; i < NV; i++) ex[i] = T(0.0);
  if ((qp[0] <= T(kEps)) && (arc_p == T(0.0))) return;
  if (qp[0] > T(kEps)) {
    T h = qp[0], u = qp[NV], v = qp[NV + 1];
    MixtT<T> m = mixt_var<NS, T, NX>(P, qp, true);
    T tg = m.red_grav + T(0.0);
    ex[1] = -(T(1.0) * m.rho_m * tg * h * Bx) + T(0.5) * m.rho_m * m.red_grav * (h * h) * T(0.0);
    ex[2] = -(T(1.0) * m.rho_m * tg * h * By) + T(0.5) * m.rho_m * m.red_grav * (h * h) * T(0.0);
    ex[3] = P.energy ? ex[1] * u + ex[2] * v : T(0.0);
  }
  if (P.radial && (arc_p > T(0.0)) && (P.h_source > 0.0)) {
    const double* tp = P.time_param;
    double t_coeff;
    if (time >= tp[3]) {
      if ((tp[2] > 0.0) && (time < tp[3] + tp[2]))
        t_coeff = 0.5 * (1.0 + cos(P.pi_g * (time - tp[3]) / tp[2]));
      else
        t_coeff = 0.0;
    } else {
      double t_rem = fmod(time, tp[0]);
      t_coeff = 0.0;
      if (tp[2] == 0.0) {
        if (t_rem <= tp[1]) t_coeff = 1.0;
      } else {
        if (t_rem < tp[2])
          t_coeff = 0.5 * (1.0 - cos(P.pi_g * t_rem / tp[2]));
        else if (t_rem <= (tp[1] - tp[2]))
          t_coeff = 1.0;
        else if (t_rem <= tp[1])
          t_coeff = 0.5 * (1.0 + cos(P.pi_g * ((t_rem - tp[1]) / tp[2] + 1.0)));
      }
    }
    if (t_coeff > 0.0) {
      double qs[NV + 2];
#pragma unroll
      for (int i = 0; i < NV + 2; i++) qs[i] = 0.0;
      qs[0] = 1.0;
      qs[3] = P.T_source;
#pragma unroll
      for (int i = 0; i < NS; i++) qs[4 + i] = P.alphas_source[i] * qs[0];
      Mixt m = mixt_var<NS, double, NX>(P, qs, true);
      const T tc = T(t_coeff);
      const T h_dot = T(P.h_source) * T(P.vel_source) * arc_p / (T(P.dx) * T(P.dy));
      const T rmx = T(m.rho_m);
      ex[0] = ex[0] + tc * h_dot * rmx;
      ex[1] = ex[1] + tc * h_dot * rmx * T(P.vel_source) * arc_x;
      ex[2] = ex[2] + tc * h_dot * rmx * T(P.vel_source) * arc_y;
      ex[3] = ex[3] + tc * h_dot * rmx * T(m.cp_mix) * T(P.T_source);
#pragma unroll
      for (int i = 0; i < NS; i++)
        ex[4 + i] = ex[4 + i] + tc * h_dot * T(P.alphas_source[i]) * T(P.rho_s[i]);
    }
  }
}

// sauter_diameter and average_density_solids of the solid volume fractions:
// src/constitutive_2d.f90:387-416
template <int NS>
__device__ inline void mean_particle(const Params& P, const double* al, double& diam,
                                     double& rho) {
  double s = 0.0, t = 0.0, r = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) s += al[i];
  if (s < kEps) {
#pragma unroll
    for (int i = 0; i < NS; i++) {
      t += P.diam_s[i];
      r += P.rho_s[i];
    }
    diam = t / (double)NS;
    rho = r / (double)NS;
    return;
  }
#pragma unroll
  for (int i = 0; i < NS; i++) {
    t += al[i] / P.diam_s[i];
    r += al[i] * P.rho_s[i];
  }
  diam = s / t;
  rho = r / s;
}

// mu(I) friction coefficient at the base for a flow of thickness h (floored
// at EPSILON), speed mod_vel and effective normal stress ens (floored at
// EPSILON), with grav_coeff = 1: src/constitutive_2d.f90:4409-4446 (model 11)
// and :4488-4525 (model 12) compute the same quantities.  Returns mu(I).
__device__ inline double mu_of_I(const Params& P, double diam, double rhop, double h,
                                 double mod_vel, double ens) {
  const double shear = 5.0 / 2.0 * mod_vel / (h * sqrt(1.0));
  const double In = diam * shear / sqrt(ens / rhop);
  return (P.mu_s * P.I_0 + P.mu_2 * In + P.muI_inf * (In * In)) / (P.I_0 + In);
}

// eval_nh_semi_impl_terms, grav_coeff = 1: src/constitutive_2d.f90:4027-4547;
// Voellmy Coulomb part (model 1, :4164-4213) or the mu(I) friction of models
// 11 and 12 (:4384-4537)
template <int NS, int NX = 0>
__device__ inline void semi_impl(const Params& P, const double* qp, double* si) {
  constexpr int NV = 4 + NS + nx_count(NX);
#pragma unroll
  for (int i = 0; i < NV; i++) si[i] = 0.0;
  double h = qp[0], u = qp[NV], v = qp[NV + 1];
  Mixt m = mixt_var<NS, double, NX>(P, qp, false);
  double w = 0.0;
  double mod_vel = sqrt(u * u + v * v + w * w);
  if (P.rheology == 1) {
    if (mod_vel > 0.0) {
      double tt = P.mu * m.rho_m * (m.red_grav + 0.0) * h * sqrt(1.0);
      tt = fmax_f(0.0, tt);
      if constexpr (has_pore(NX))  // :4191-4198
        tt = fmax_f(0.0, tt - P.mu * sqrt(1.0) * qp[pore_at<NS, NX>()]);
      tt = fmax_f(0.0, tt);
      si[1] = si[1] - tt * u / mod_vel;
      si[2] = si[2] - tt * v / mod_vel;
    }
  } else if (mod_vel > 0.0) {
    double diam, rhop;
    if (NS > 1) {
      double al[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) al[i] = qp[4 + i] / qp[0];
      mean_particle<NS>(P, al, diam, rhop);
    } else {
      diam = P.diam_s[0];
      rhop = P.rho_s[0];
    }
    if (h < kEps) h = kEps;
    double vse = m.rho_m * (m.red_grav + 0.0) * h;
    if constexpr (has_pore(NX))  // :4423-4427
      vse = m.rho_m * (m.red_grav + 0.0) * h - qp[pore_at<NS, NX>()];
    vse = fmax_f(vse, 0.0);
    double ens = fmax_f(0.0, vse * sqrt(1.0));
    if (ens < kEps) ens = kEps;
    const double tt = mu_of_I(P, diam, rhop, h, mod_vel, ens) * ens;
    si[1] = si[1] - tt * (u / mod_vel);
    si[2] = si[2] - tt * (v / mod_vel);
  }
}

// xi_temp of eval_implicit_terms for a cell with stochastic value Z: xi + Z
// limited to [1, 1e5] with the stochastic friction, xi without
// (src/constitutive_2d.f90:3663-3673)
__device__ __forceinline__ double cell_xi(const Params& P, const double* Z, int c) {
  if (!P.stochastic) return P.xi;
  double x = P.xi + Z[c];
  if (x < 1.0) x = 1.0;
  if (x > 1e5) x = 1e5;
  return x;
}

// Carman-Kozeny grain size d * psi of the dynamic permeability: the class's
// diameter and sphericity for one class; the Sauter diameter and the
// surface-area-weighted sphericity for several (src/constitutive_2d.f90:
// 3853-3868, :4978-4990)
template <int NS>
__device__ inline double perm_grain(const Params& P, const double* al) {
  if (NS > 1) {
    double s = 0.0, t = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) s += al[i];
    double diam;
    if (s < kEps) {
#pragma unroll
      for (int i = 0; i < NS; i++) t += P.diam_s[i];
      diam = t / (double)NS;
    } else {
#pragma unroll
      for (int i = 0; i < NS; i++) t += al[i] / P.diam_s[i];
      diam = s / t;
    }
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) a += al[i] * (P.diam_s[i] * P.diam_s[i]) * P.sph_s[i];
#pragma unroll
    for (int i = 0; i < NS; i++) b += al[i] * (P.diam_s[i] * P.diam_s[i]);
    return diam * (a / b);
  }
  return P.diam_s[0] * P.sph_s[0];
}

// 10_wp**(-0.28_wp) as the reference build folds it at compile time
constexpr double kTenPowM028 = 0x1.0cb3903cf06e8p-1;

// 1 - sum_n pascal(n) tt**(N + n + 1), tt = (x - a_t)/(max_pack - a_t)
// clamped to [0, 1]: f_inhibit's generalised smoothstep (STATIC, DYNAMIC;
// src/constitutive_2d.f90:3904-3927, :4911-4934)
__device__ inline double finh_sum(const Params& P, double x, double at) {
  const double width = P.max_pack - at;
  double tt = (x - at) / width;
  if (tt < 0.0) tt = 0.0;
  if (tt > 1.0) tt = 1.0;
  double s = 0.0;
  for (int n = 0; n <= P.N_inh; n++) {
    const double term = P.pascal[n] * pow_cr(tt, (double)(P.N_inh + n) + 1.0);
    s = s + term;
  }
  return 1.0 - s;
}

// the pore-pressure row of eval_implicit_terms in complex arithmetic:
// src/constitutive_2d.f90:3844-3993
template <int NS>
__device__ inline cx pore_source_c(const Params& P, cx h, cx T, cx rho_m, const cx* alphas,
                                   cx exc) {
  const double h_threshold = 1.0e-10;
  const double gamma_gas = P.cp_a / (P.cp_a - P.R_a);
  const cx gas_compr = C(1.0 / (gamma_gas * P.pres), 0.0);
  const cx rho_gas = cdiv(C(P.pres, 0.0), cmul(C(P.R_a, 0.0), T));
  cx sa = C(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) sa = cadd(sa, alphas[i]);
  const cx porosity = csub(C(1.0, 0.0), sa);
  double hyd = P.hyd_perm;
  if (P.dyn_perm) {
    double al[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) al[i] = alphas[i].re;
    const double ds = perm_grain<NS>(P, al);
    const cx num = cmul(cpow_c(porosity, 3.0), C(ds * ds, 0.0));
    const cx den = cmul(C(150.0, 0.0), cpow_c(csub(C(1.0, 0.0), porosity), 2.0));
    hyd = cdiv(num, den).re;
  }
  const cx D = cdiv(C(hyd, 0.0),
                    cmul(cmul(cmul(porosity, C(P.kin_visc_c, 0.0)), rho_gas), gas_compr));
  if (!((exc.re > 0.0) && (h.re > 0.0))) return C(0.0, 0.0);
  cx f;
  if (P.finh == 0) {
    f = C(1.0, 0.0);
  } else if (P.finh == 1) {
    f = (sa.re < P.max_pack) ? C(1.0, 0.0) : C(0.0, 0.0);
  } else {
    double at = P.alpha_trans;
    if (P.finh == 3) {
      at = cmul(C(kTenPowM028, 0.0), cpow_c(h, 0.12)).re;
      if (at >= P.max_pack) at = 0.99 * P.max_pack;
    }
    const double width = P.max_pack - at;
    double tt = (sa.re - at) / width;
    if (tt < 0.0) tt = 0.0;
    if (tt > 1.0) tt = 1.0;
    cx sc = C(0.0, 0.0);
    for (int n = 0; n <= P.N_inh; n++) {
      const double term = P.pascal[n] * pow_cr(tt, (double)(P.N_inh + n) + 1.0);
      sc = cadd(sc, C(term, 0.0));
    }
    f = csub(C(1.0, 0.0), sc);
  }
  const double c2 = (P.pi_g / 2.0) * (P.pi_g / 2.0);
  const cx hm = (h_threshold > h.re) ? C(h_threshold, 0.0) : h;
  const cx r = cmul(cmul(cdiv(cmul(cmul(rho_m, C(c2, 0.0)), D), hm), exc), f);
  return C(-r.re, -r.im);
}

// c_phys_var + eval_implicit_terms (Voellmy turbulent part) in complex
// arithmetic: src/constitutive_2d.f90:1154-1322, :3513-3683, :3994-4009.
// Only the momentum rows are non-zero; they are returned in nh1, nh2.
template <int NS, int NX = 0>
__device__ inline void implicit_c(const Params& P, const cx* qj, cx& nh1, cx& nh2, double xi,
                                  cx* nhp = nullptr) {
  nh1 = C(0.0, 0.0);
  nh2 = C(0.0, 0.0);
  // mu(I) model 11 has no implicit friction term (:3660 takes models 1, 12)
  if (!has_pore(NX) && P.rheology == 11) return;
  cx u, v, rho_m, T, hh, exc, alphas[NS];
  if (qj[0].re > kEps) {
    cx inv_cqj1 = cdiv(C(1.0, 0.0), qj[0]);
    exc = has_pore(NX) ? cmul(qj[pore_at<NS, NX>()], inv_cqj1) : C(0.0, 0.0);  // :1195
    cx xs[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) xs[i] = cmul(qj[4 + i], inv_cqj1);
    cx xs_tot = C(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) xs_tot = cadd(xs_tot, xs[i]);
    cx xc = csub(C(1.0, 0.0), xs_tot);
    cx sp_heat_c = cdiv(cadd(cmul(csub(xc, C(0.0, 0.0)), C(P.cp_a, 0.0)), C(0.0, 0.0)), xc);
    cx dot = C(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) dot = cadd(dot, cmul(cconj(xs[i]), C(P.cp_s[i], 0.0)));
    cx sp_heat_mix = cadd(dot, cmul(xc, sp_heat_c));
    if (qj[0].re > P.eps_sing) {
      if (P.energy) {
        cx ke = cmul(cmul(C(0.5, 0.0), cadd(csq(qj[1]), csq(qj[2]))), inv_cqj1);
        T = cdiv(csub(qj[3], ke), cmul(qj[0], sp_heat_mix));
      } else {
        T = cdiv(qj[3], cmul(qj[0], sp_heat_mix));
      }
      if (T.re <= 0.0) T = C(P.T_amb, 0.0);
    } else {
      T = C(P.T_amb, 0.0);
    }
    cx Rc = cdiv(cadd(cmul(csub(xc, C(0.0, 0.0)), C(P.R_a, 0.0)), C(0.0, 0.0)), xc);
    cx inv_rho_c = cmul(cmul(Rc, T), C(P.inv_pres, 0.0));
    cx d2 = C(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) d2 = cadd(d2, cmul(cconj(xs[i]), C(P.inv_rho_s[i], 0.0)));
    cx inv_rhom = cadd(d2, cmul(xc, inv_rho_c));
    rho_m = cdiv(C(1.0, 0.0), inv_rhom);
    if constexpr (has_pore(NX)) {  // :1286-1290
#pragma unroll
      for (int i = 0; i < NS; i++) alphas[i] = cmul(cmul(rho_m, xs[i]), C(P.inv_rho_s[i], 0.0));
      hh = cmul(qj[0], inv_rhom);
    }
    if (qj[0].re > P.eps_sing) {
      u = cmul(qj[1], inv_cqj1);
      v = cmul(qj[2], inv_cqj1);
    } else {
      cx c4 = cmul(qj[0], cmul(qj[0], cmul(qj[0], qj[0])));
      cx den = csqrt_c(cadd(c4, C(P.eps_sing4, 0.0)));
      cx s2 = C(kSqrt2, 0.0);
      u = cdiv(cmul(cmul(s2, qj[0]), qj[1]), den);
      v = cdiv(cmul(cmul(s2, qj[0]), qj[2]), den);
    }
  } else {
    u = C(0.0, 0.0);
    v = C(0.0, 0.0);
    rho_m = C(P.rho_a_amb, 0.0);
    T = C(P.T_amb, 0.0);
    hh = C(0.0, 0.0);
    exc = C(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) alphas[i] = C(0.0, 0.0);
  }
  if (P.rheology != 11) {
    cx red_grav = cmul(cdiv(csub(rho_m, C(P.rho_a_amb, 0.0)), rho_m), C(P.grav, 0.0));
    cx w = C(0.0, 0.0);
    cx mod_vel2 = cadd(cadd(csq(u), csq(v)), csq(w));
    cx mod_vel = csqrt_c(mod_vel2);
    if (mod_vel.re != 0.0) {
      cx ts = cmul(cdiv(cmul(rho_m, red_grav), C(xi, 0.0)), mod_vel2);
      nh1 = csub(nh1, cmul(ts, cdiv(u, mod_vel)));
      nh2 = csub(nh2, cmul(ts, cdiv(v, mod_vel)));
    }
  }
  if constexpr (has_pore(NX)) *nhp = pore_source_c<NS>(P, hh, T, rho_m, alphas, exc);
}

// The residual evaluation of the Newton solve calls the implicit terms at a
// real state: the reference converts r_qj to complex (imaginary parts +0) and
// runs the complex code (src/constitutive_2d.f90:3513-3683).  With imaginary
// parts that are all zeros, each complex operation's real part is the real
// operation on the real parts plus or minus a signed zero:
//   (a + 0i)(b + 0i)  real part  a*b - (+-0)            = a*b
//   (a + 0i)/(c + 0i) (Annex G)  scalbn((a*c')/(c'*c'), -k),  k = logb|c|,
//                                c' = scalbn(c, -k)
//   sqrt(x + 0i), x > 0          sqrt(x)
// and every imaginary part stays a signed zero, provided no real
// intermediate is zero (the signed zeros could then decide the sign of a
// zero), infinite or NaN (a NaN imaginary part would reach the real parts).
// implicit_r evaluates the real parts that way and reports whether that
// proviso held for every intermediate; when it did not, the caller takes the
// complex path.  The result is then bit for bit the complex path's real part.
__device__ __forceinline__ bool nz_finite(double v) {
  return (fabs(v) > 0.0) & (fabs(v) < INFINITY);
}
// real part of cdiv(C(a, +-0), C(c, +-0)); ok cleared unless a, c and the
// quotient are non-zero and finite
__device__ __forceinline__ double cdiv_re(double a, double c, bool& ok) {
  ok = ok & nz_finite(a) & nz_finite(c);
  const int k = ilogb(c);  // logb(fmax(|c|, 0)) for a finite non-zero c
  const double cs = scalbn(c, -k);
  const double x = scalbn((a * cs) / (cs * cs), -k);
  ok = ok & nz_finite(x);
  return x;
}
template <int NS>
__device__ inline bool implicit_r(const Params& P, const double* q, double& nh1, double& nh2,
                                  double xi) {
  nh1 = 0.0;
  nh2 = 0.0;
  if (P.rheology == 11) return true;  // no implicit friction term
  // dry: u = v = 0 and mod_vel = 0 on the complex path, both terms +0
  if (!(q[0] > kEps)) return true;
  if (!(q[0] > P.eps_sing)) return false;  // thin-layer velocity regularisation: complex path
  bool ok = nz_finite(q[0]);
  const double inv = cdiv_re(1.0, q[0], ok);
  double xs[NS];
  double xs_tot = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    xs[i] = q[4 + i] * inv;
    ok = ok & nz_finite(xs[i]);
    xs_tot = xs_tot + xs[i];
  }
  const double xc = 1.0 - xs_tot;
  double t = (xc - 0.0) * P.cp_a;
  ok = ok & nz_finite(t);
  const double sp_heat_c = cdiv_re(t + 0.0, xc, ok);
  double dot = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double m = xs[i] * P.cp_s[i];
    ok = ok & nz_finite(m);
    dot = dot + m;
  }
  const double m2 = xc * sp_heat_c;
  ok = ok & nz_finite(m2);
  const double sp_heat_mix = dot + m2;
  const double den = q[0] * sp_heat_mix;
  ok = ok & nz_finite(den);
  double T;
  if (P.energy) {
    const double a1 = q[1] * q[1], a2 = q[2] * q[2];
    ok = ok & nz_finite(a1) & nz_finite(a2);
    double ke = 0.5 * (a1 + a2);
    ok = ok & nz_finite(ke);
    ke = ke * inv;
    ok = ok & nz_finite(ke);
    T = cdiv_re(q[3] - ke, den, ok);
  } else {
    T = cdiv_re(q[3], den, ok);
  }
  if (T <= 0.0) T = P.T_amb;
  t = (xc - 0.0) * P.R_a;
  ok = ok & nz_finite(t);
  const double Rc = cdiv_re(t + 0.0, xc, ok);
  double irc = Rc * T;
  ok = ok & nz_finite(irc);
  irc = irc * P.inv_pres;
  ok = ok & nz_finite(irc);
  double d2 = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double m = xs[i] * P.inv_rho_s[i];
    ok = ok & nz_finite(m);
    d2 = d2 + m;
  }
  const double m3 = xc * irc;
  ok = ok & nz_finite(m3);
  const double rho_m = cdiv_re(1.0, d2 + m3, ok);
  const double u = q[1] * inv, v = q[2] * inv;
  ok = ok & nz_finite(u) & nz_finite(v);
  const double red_grav = cdiv_re(rho_m - P.rho_a_amb, rho_m, ok) * P.grav;
  ok = ok & nz_finite(red_grav);
  const double uu = u * u, vv = v * v;
  ok = ok & nz_finite(uu) & nz_finite(vv);
  const double mod_vel2 = (uu + vv) + 0.0;
  const double mod_vel = sqrt(mod_vel2);
  ok = ok & nz_finite(mod_vel);
  const double rg = rho_m * red_grav;
  ok = ok & nz_finite(rg);
  const double ts = cdiv_re(rg, xi, ok) * mod_vel2;
  ok = ok & nz_finite(ts);
  const double p1 = ts * cdiv_re(u, mod_vel, ok), p2 = ts * cdiv_re(v, mod_vel, ok);
  ok = ok & nz_finite(p1) & nz_finite(p2);
  nh1 = 0.0 - p1;
  nh2 = 0.0 - p2;
  return ok;
}

// settling_velocity (Schiller-Naumann): src/constitutive_2d.f90:5326-5404
__device__ inline double settling_velocity(const Params& P, double diam, double rhos,
                                           double rhoc, double ikv, const ddm::dd* tab,
                                           PhaseClk* pc = nullptr) {
  double inv_sqrt_C_D = 1.0;
  if (rhos <= rhoc) return 0.0;
  double const_part = sqrt(1.3333333333333333 * (rhos / rhoc - 1.0) * diam * P.grav);
  double sv = const_part * inv_sqrt_C_D;
  double Rey = diam * sv * ikv;
  if (Rey <= 1000.0) {
#pragma unroll 1
    for (int i = 0; i < 20; i++) {
      pc_add(pc, PC_SETTLE_ITS, 1ULL);
      double old = sv;
      inv_sqrt_C_D = sqrt(Rey / (24.0 * (1.0 + 0.15 * pow_cr(Rey, 0.687, tab))));
      sv = const_part * inv_sqrt_C_D;
      if (fabs(old - sv) / old < 1.0e-6) return sv;
      Rey = diam * sv * ikv;
    }
  }
  return sv;
}

// eval_mass_exchange_terms without pore pressure / vaporisation:
// src/constitutive_2d.f90:4574-5153 (erosion :4731-4758).  erod: the cell's
// erodible layer per class; ers receives the erosion rates.  Returns true
// when the reference would pause on a negative deposition rate.
// LANES = 2: the two lanes of an aligned pair evaluate the same cell and
// must call this together; lane `lane` computes the settling velocities of
// the classes i with i % 2 == lane and the pair swaps them, so the slowest
// part (the drag fixed point) runs on both lanes at once.
template <int NS, int LANES = 1, int NX = 0>
__device__ inline bool mass_exchange(const Params& P, const double* qp, const double* erod,
                                     double dt, double* dep, double* eq, double* ers,
                                     double& topo, const ddm::dd* tab, int lane = 0,
                                     PhaseClk* pc = nullptr) {
  constexpr int NV = 4 + NS + nx_count(NX);
  topo = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    ers[i] = 0.0;
    dep[i] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < NV; i++) eq[i] = 0.0;
  if (qp[0] <= kEps) return false;
  double h = qp[0], u = qp[NV], v = qp[NV + 1];
  double alphas[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) alphas[i] = qp[4 + i] / qp[0];
  double at = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) at += alphas[i];
  double w = 0.0;
  double mod_vel2 = u * u + v * v + w * w;
  if (P.erosion_coeff > 0.0) {
    // Fagents & Baloga (2006) with the solid volume fraction
    const double tot_erosion = P.erosion_coeff * sqrt(mod_vel2) * h * (1.0 - at);
    const double tse0 = tot_erosion * (1.0 - P.erodible_porosity);
#pragma unroll
    for (int i = 0; i < NS; i++) ers[i] = P.erodible_fract[i] * tse0;
  }
#pragma unroll
  for (int i = 0; i < NS; i++) ers[i] = fmax_f(0.0, fmin_f(ers[i], erod[i] / dt));
  double tse = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) tse += ers[i];
  // tot_erosion = tse / (1 - porosity): exactly tse without erosion (porosity 0)
  const double te = (P.erodible_porosity == 0.0) ? tse : tse / (1.0 - P.erodible_porosity);
  double cpe = te * P.erodible_porosity;
  if (at <= 0.0) return false;
  double T = qp[3];
  Mixt m = mixt_var<NS, double, NX>(P, qp, true);
  double ikv = 1.0 / P.kin_visc_c;
  bool err = false;
  // settling velocities (pure functions of the class and rho_c) and the
  // hindered-settling factor, the same for every class of the cell
  double sv[NS];
  bool any = false;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    sv[i] = 0.0;
    any = any || ((alphas[i] > 0.0) && P.settling);
  }
  if constexpr (LANES == 1) {
#pragma unroll
    for (int i = 0; i < NS; i++)
      if ((alphas[i] > 0.0) && P.settling)
        sv[i] = settling_velocity(P, P.diam_s[i], P.rho_s[i], m.rho_c, ikv, tab);
  } else {
    // classes i0 and i0 + 1 go through ONE call on the pair (each lane with
    // its own class's constants), so the two fixed-point iterations run side
    // by side rather than one after the other with half the lanes masked
#pragma unroll
    for (int i0 = 0; i0 < NS; i0 += 2) {
      const bool two = i0 + 1 < NS;
      const bool mine = (lane == 0) || two;
      const double a = (lane == 0 || !two) ? alphas[i0] : alphas[two ? i0 + 1 : i0];
      const double d = (lane == 0 || !two) ? P.diam_s[i0] : P.diam_s[two ? i0 + 1 : i0];
      const double r = (lane == 0 || !two) ? P.rho_s[i0] : P.rho_s[two ? i0 + 1 : i0];
      double s = 0.0;
      const unsigned long long c0 = pc_now();
      if (mine && (a > 0.0) && P.settling) s = settling_velocity(P, d, r, m.rho_c, ikv, tab, pc);
      pc_add(pc, PC_SETTLE, pc_now() - c0);
      const double o = __shfl_xor(s, 1);
      sv[i0] = (lane == 0) ? s : o;
      if (two) sv[i0 + 1] = (lane == 0) ? o : s;
    }
  }
  // the two remaining powers of the cell, the hindered-settling factor
  // (1 - a/0.6)**4.65 and the entrainment's Ri**2.4, are independent: with
  // a lane pair, lane 0 takes the first and lane 1 the second
  const bool ent = P.entrainment && (h > 0.0) && (m.Ri > 0.0);
  double hindered, pent;
  const unsigned long long c1 = pc_now();
  if constexpr (LANES == 2) {
    const bool mine = (lane == 0) ? any : ent;
    const double px = (lane == 0) ? 1.0 - fmin_f(1.0, at / 0.6) : m.Ri;
    const double pr = mine ? pow_cr(px, (lane == 0) ? 4.65 : 2.4, tab) : 0.0;
    const double po = __shfl_xor(pr, 1);
    hindered = (lane == 0) ? pr : po;
    pent = (lane == 0) ? po : pr;
  } else {
    hindered = any ? pow_cr(1.0 - fmin_f(1.0, at / 0.6), 4.65, tab) : 0.0;
    pent = ent ? pow_cr(m.Ri, 2.4, tab) : 0.0;
  }
  pc_add(pc, PC_POW, pc_now() - c1);
#pragma unroll
  for (int i = 0; i < NS; i++) {
    if ((alphas[i] > 0.0) && P.settling) {
      dep[i] = alphas[i] * sv[i] * 1.0;
      dep[i] = dep[i] * hindered;
      dep[i] = fmin_f(dep[i], h * alphas[i] / dt);
      if (dep[i] < 0.0) err = true;
    }
  }
  double sdep = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) sdep += dep[i];
  double cpl = 0.0 + P.coeff_porosity * sdep;
  double max_pack = 0.64;
  if constexpr (has_pore(NX)) {
    max_pack = P.max_pack;
    // gas loss driven by the excess pore pressure, inhibited by f_inhibit:
    // src/constitutive_2d.f90:4887-5032
    if (P.gas_loss && at < P.max_pack) {
      const double pexc = qp[pore_at<NS, NX>()];
      double f_inh = 1.0;
      if (P.finh == 1) {
        f_inh = (at < P.max_pack) ? 1.0 : 0.0;
      } else if (P.finh == 2) {
        f_inh = finh_sum(P, at, P.alpha_trans);
      } else if (P.finh == 3) {
        if (h < kEps) h = kEps;
        double atd = kTenPowM028 * pow_cr(h, 0.12, tab);
        if (atd >= P.max_pack) atd = 0.99 * P.max_pack;
        f_inh = finh_sum(P, at, atd);
      }
      double hyd = P.hyd_perm;
      if (P.dyn_perm) {
        const double ds = perm_grain<NS>(P, alphas), pr = 1.0 - at;
        hyd = ((pr * pr) * pr * (ds * ds)) / (150.0 * (at * at));
      }
      // MAX(1.e-5, r_h) with a default-real literal
      const double hmin = (double)1.0e-5f;
      const double vel = hyd / (P.kin_visc_c * m.rho_c) / fmax_f(hmin, h) * 0.5 * P.pi_g * pexc;
      cpl = cpl + vel * f_inh;
    }
  }
  // the continuous-phase loss is limited by the maximum solid packing (0.64
  // without pore pressure) and by the continuous phase present
  // (src/constitutive_2d.f90:5034-5041)
  cpl = fmin_f(cpl, h * fmax_f(0.0, max_pack - at) / dt);
  cpl = fmin_f(cpl, (h * (1.0 - at) - P.coeff_porosity * (h * at - dt * sdep)) / dt);
  // with erodible_deposit_flag the reference sets T_erodible to the cell's T
  // before using it (src/constitutive_2d.f90:5044)
  const double T_er = P.erodible_deposit ? T : P.T_erodible;
  double air = 0.0;
  if (ent) {
    double ec = 0.075 / sqrt(1.0 + 718.0 * pent);
    air = ec * sqrt(mod_vel2);
  }
  double rdt = 0.0, rert = 0.0, A = 0.0, D = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) rdt += P.rho_s[i] * dep[i];
#pragma unroll
  for (int i = 0; i < NS; i++) rert += P.rho_s[i] * ers[i];
#pragma unroll
  for (int i = 0; i < NS; i++) A += P.rho_s[i] * P.cp_s[i] * dep[i];
#pragma unroll
  for (int i = 0; i < NS; i++) D += P.rho_s[i] * P.cp_s[i] * ers[i];
  double rc = m.rho_c, cpc = m.cp_c;
  eq[0] = P.rho_a_amb * air + rert - rdt + P.rho_c_sub * cpe - rc * cpl;
  eq[1] = -(u * (rdt + rc * cpl));
  eq[2] = -(v * (rdt + rc * cpl));
  if (P.energy)
    eq[3] = -(T * (A + rc * cpc * cpl)) - 0.5 * mod_vel2 * (rdt + rc * cpl) +
            T_er * (D + P.rho_c_sub * cpc * cpe) + P.T_amb * P.cp_a * P.rho_a_amb * air;
  else
    eq[3] = -(T * (A + rc * cpc * cpl)) + T_er * (D + P.rho_c_sub * cpc * cpe) +
            P.T_amb * P.cp_a * P.rho_a_amb * air;
#pragma unroll
  for (int i = 0; i < NS; i++) eq[4 + i] = P.rho_s[i] * (ers[i] - dep[i]);
  if constexpr (has_stoch(NX)) eq[stoch_at<NS>()] = eq[0] * qp[stoch_at<NS>()];  // :5132-5137
  if constexpr (has_pore(NX)) eq[pore_at<NS, NX>()] = eq[0] * qp[pore_at<NS, NX>()];  // :5140-5145
  // rate of change of the topography (src/constitutive_2d.f90:5149), used
  // only with topography change
  if (P.topo_change) {
    double dep_tot = 0.0, ers_tot = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) dep_tot += dep[i];
#pragma unroll
    for (int i = 0; i < NS; i++) ers_tot += ers[i];
    topo = (dep_tot - ers_tot) / (1.0 - P.erodible_porosity);
  }
  return err;
}

// x / |x| as the correctly rounded quotient gives it
template <typename T>
__device__ __forceinline__ T sign_quot(T x) {
  return (fabs(x) > T(0.0) && fabs(x) < T(INFINITY)) ? copysign(T(1.0), x) : T(__builtin_nan(""));
}

// limit with the minmod limiter: src/geometry_2d.f90:1776-1806, :1855-1876
template <typename T>
__device__ __forceinline__ T limit_minmod(T v0, T v1, T v2, T z0, T z1, T z2) {
  T a = (v2 - v1) / (z2 - z1);
  T b = (v1 - v0) / (z1 - z0);
  T m = fmin_f(fabs(a), fabs(b));
  if (m <= T(0.0)) return T(0.0);
  // a / |a| without the division: exactly +-1 for a finite non-zero a and
  // NaN for 0, +-Inf and NaN (0/0, Inf/Inf, NaN), as the quotient is
  T sa = sign_quot(a), sb = sign_quot(b);
  return T(0.5) * (sa + sb) * m;
}

}  // namespace imex
