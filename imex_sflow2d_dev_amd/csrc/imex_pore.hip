// The time step of runs with pore pressure (pore_pressure_flag) or the
// transported noise (stoch_transport_flag) on gfx950: the general path.
//
// The transported noise adds the conserved variable q1 * Z after the solid
// classes (idx_stoch, before the pore pressure): a passive tracer advected
// with the flow (src/constitutive_2d.f90:2428-2431, :2490-2493), an explicit
// row of the Newton system, and the source eqns_term(1) * Z of the mass
// exchange (:5132-5137); the friction keeps reading the fixed noise plane.
//
// Pore pressure adds a conserved variable, q1 * p_exc, after the solid
// classes (n_vars = 5 + n_solid), a third implicit equation (its decay by
// diffusion through the bed, with f_inhibit and the Carman-Kozeny dynamic
// permeability, src/constitutive_2d.f90:3844-3993), so the Newton system
// of each cell is 3 x 3 and the reference solves it with DGESV
// (src/solver_2d.f90:1905-1917), the effective normal stress of the
// friction laws (:4191-4198, :4423-4427, :4502-4506) and the gas loss of the
// mass exchange (:4887-5032).  The decks that use it (TESTS/TEST_DAMBREAK)
// are small grids (300 x 15), so this path is written for clarity and
// parity rather than for the last microsecond: one thread per cell or per
// face over the whole grid, a launch per phase of the reference's step:
//   k_pp_prim   primitive states of every cell (dry ones: zeros, T_ambient)
//   k_pp_recon  MUSCL/minmod face states of every cell (the reference's
//               reconstruction; cells outside the active set give the dry
//               states a fresh evaluation gives them)
//   k_pp_faces  local speeds and Kurganov-Tadmor fluxes of every face, with
//               the solid-flux caps and the zero-velocity rule
//   k_pp_cells  active cells: stage-1 right-hand side divF - expl, CFL
//               candidate (one atomicMin per cell on the bit pattern)
//   k_pp_update active cells: the step's dt, stage 2 (semi-implicit friction,
//               Newton with the 3 x 3 solve), assembly checks, mass exchange
//               with the gas loss, the driver epilogue
// with the active set of k_mask (every tile each step) and the dt history
// committed by the next launch, as in the other device-driven steps.  The
// physics functions are the shared ones of imex_physics.h instantiated for
// one extra variable (NX = 1).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "imex_handle.h"

namespace {

using namespace imex;

constexpr int PTHR = 256;

struct PoreBufs {
  double* qp;                 // [NP][N] primitive states
  double *qpL, *qpR;          // [NP][NFX] x-face sides (L: of the cell west of the face)
  double *qpB, *qpT;          // [NP][NFY] y-face sides (B: of the cell south of the face)
  double *Hx, *Hy;            // [NV][NFX], [NV][NFY] face fluxes
  double *amx, *amy;          // [NFX], [NFY] max(a+, -a-) of each face
};

__device__ __forceinline__ bool step_done(const Grid& G) { return G.st->done != 0u; }

// primitive state of one cell: qc_to_qp when wet, else zeros and T_ambient
// (src/solver_2d.f90:944-957 for timestep, :1300-1308 for the stages)
template <int NS, int NX>
__global__ __launch_bounds__(PTHR) void k_pp_prim(Params P, Grid G, PoreBufs B) {
  constexpr int NV = 4 + NS + nx_count(NX), NP = NV + 2;
  if (step_done(G)) return;
  const int N = G.N;
  for (int c = blockIdx.x * PTHR + threadIdx.x; c < N; c += gridDim.x * PTHR) {
    double q[NV], qp[NP], pd;
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = G.q[(size_t)i * N + c];
    if (q[0] > 0.0) {
      qc_to_qp<NS, double, NX>(P, q, qp, pd);
    } else {
#pragma unroll
      for (int i = 0; i < NP; i++) qp[i] = 0.0;
      qp[3] = P.T_amb;
    }
#pragma unroll
    for (int i = 0; i < NP; i++) B.qp[(size_t)i * N + c] = qp[i];
  }
}

__device__ __forceinline__ void qp2_uv(const double* qp, double& u, double& v) {
  // qp_to_qp2 (src/constitutive_2d.f90:2164-2188): u, v of a face state
  if (qp[0] <= 0.0) {
    u = 0.0;
    v = 0.0;
  } else {
    u = qp[1] / qp[0];
    v = qp[2] / qp[0];
  }
}

__device__ __forceinline__ double minmod3(double v0, double v1, double v2, double z0, double z1,
                                          double z2) {
  return limit_minmod<double>(v0, v1, v2, z0, z1, z2);
}

// reconstruction of one cell (src/solver_2d.f90:3095-3924, minmod limiter,
// boundary flags 1 with value 0, no lateral source): its W, E, S and N face
// states into the face arrays, the boundary faces' outer sides copied from
// the inner ones (:3926-3935)
template <int NS, int NX>
__global__ __launch_bounds__(PTHR) void k_pp_recon(Params P, Grid G, PoreBufs B) {
  constexpr int NV = 4 + NS + nx_count(NX), NP = NV + 2;
  if (step_done(G)) return;
  const int nx = G.nx, ny = G.ny, N = G.N;
  const size_t NFX = (size_t)(nx + 1) * ny, NFY = (size_t)nx * (ny + 1);
  for (int c = blockIdx.x * PTHR + threadIdx.x; c < N; c += gridDim.x * PTHR) {
    const int j = c % nx, k = c / nx;
    auto at = [&](int cc, int i) { return B.qp[(size_t)i * N + cc]; };
    double W[NP], E[NP], S[NP], Nn[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) W[i] = E[i] = S[i] = Nn[i] = at(c, i);
    const double xs1 = G.xc[j], ys1 = G.yc[k];
    double xs0, xs2, ys0, ys2;
    if (j == 0) {
      xs0 = G.xs[0];
      xs2 = G.xc[1];
    } else if (j == nx - 1) {
      xs0 = G.xc[nx - 2];
      xs2 = G.xs[nx];
    } else {
      xs0 = G.xc[j - 1];
      xs2 = G.xc[j + 1];
    }
    if (k == 0) {
      ys0 = G.ys[0];
      ys2 = G.yc[1];
    } else if (k == ny - 1) {
      ys0 = G.yc[ny - 2];
      ys2 = G.ys[ny];
    } else {
      ys0 = G.yc[k - 1];
      ys2 = G.yc[k + 1];
    }
    // boundary slopes (:3190-3290, :3362-3420): flag 0 limits against the
    // Dirichlet value at the boundary face, 1 takes the value as the slope,
    // 2 the one-sided difference; the boundary face takes a Dirichlet value
    // (:3316-3356, :3452-3482)
#pragma unroll
    for (int i = 0; i < NV; i++) {
      const double v1 = at(c, i);
      double pr;
      if (j == 0) {
        const int fl = P.bcf[0][i];
        pr = (fl == 0) ? minmod3(P.bcv[0][i], v1, at(c + 1, i), xs0, xs1, xs2)
                       : ((fl == 1) ? P.bcv[0][i] : (at(c + 1, i) - v1) * P.one_by_dx);
      } else if (j == nx - 1) {
        const int fl = P.bcf[1][i];
        pr = (fl == 0) ? minmod3(at(c - 1, i), v1, P.bcv[1][i], xs0, xs1, xs2)
                       : ((fl == 1) ? P.bcv[1][i] : (v1 - at(c - 1, i)) * P.one_by_dx);
      } else {
        pr = minmod3(at(c - 1, i), v1, at(c + 1, i), xs0, xs1, xs2);
      }
      double dq = 1.0 * P.dx2 * pr;
      W[i] = v1 - dq;
      E[i] = v1 + dq;
      if (j == 0) {
        if (P.bcf[0][i] == 0) W[i] = P.bcv[0][i];
        else if (i == 1) W[i] = fmin_f(W[i], 0.0);
      }
      if (j == nx - 1) {
        if (P.bcf[1][i] == 0) E[i] = P.bcv[1][i];
        else if (i == 1) E[i] = fmax_f(E[i], 0.0);
      }
      if (k == 0) {
        const int fl = P.bcf[2][i];
        pr = (fl == 0) ? minmod3(P.bcv[2][i], v1, at(c + nx, i), ys0, ys1, ys2)
                       : ((fl == 1) ? P.bcv[2][i] : (at(c + nx, i) - v1) * P.one_by_dy);
      } else if (k == ny - 1) {
        const int fl = P.bcf[3][i];
        pr = (fl == 0) ? minmod3(at(c - nx, i), v1, P.bcv[3][i], ys0, ys1, ys2)
                       : ((fl == 1) ? P.bcv[3][i] : (v1 - at(c - nx, i)) * P.one_by_dy);
      } else {
        pr = minmod3(at(c - nx, i), v1, at(c + nx, i), ys0, ys1, ys2);
      }
      dq = 1.0 * P.dy2 * pr;
      S[i] = v1 - dq;
      Nn[i] = v1 + dq;
      if (k == 0) {
        if (P.bcf[2][i] == 0) S[i] = P.bcv[2][i];
        else if (i == 2) S[i] = fmin_f(S[i], 0.0);
      }
      if (k == ny - 1) {
        if (P.bcf[3][i] == 0) Nn[i] = P.bcv[3][i];
        else if (i == 2) Nn[i] = fmax_f(Nn[i], 0.0);
      }
    }
    // u and v with the limiters of hu and hv; at a boundary the outer
    // stencil point is the face state's own velocity
#pragma unroll
    for (int ii = 0; ii < 2; ii++) {
      const int i = NV + ii;
      const double v1 = at(c, i);
      double u2, w2, s0, s2;
      if (j == 0) {
        qp2_uv(W, u2, w2);
        s0 = ii ? w2 : u2;
        s2 = at(c + 1, i);
      } else if (j == nx - 1) {
        qp2_uv(E, u2, w2);
        s0 = at(c - 1, i);
        s2 = ii ? w2 : u2;
      } else {
        s0 = at(c - 1, i);
        s2 = at(c + 1, i);
      }
      double dq = 1.0 * P.dx2 * minmod3(s0, v1, s2, xs0, xs1, xs2);
      W[i] = v1 - dq;
      E[i] = v1 + dq;
      if (j == 0) {
        qp2_uv(W, u2, w2);
        W[i] = ii ? w2 : u2;
      } else if (j == nx - 1) {
        qp2_uv(E, u2, w2);
        E[i] = ii ? w2 : u2;
      }
      if (k == 0) {
        qp2_uv(S, u2, w2);
        s0 = ii ? w2 : u2;
        s2 = at(c + nx, i);
      } else if (k == ny - 1) {
        qp2_uv(Nn, u2, w2);
        s0 = at(c - nx, i);
        s2 = ii ? w2 : u2;
      } else {
        s0 = at(c - nx, i);
        s2 = at(c + nx, i);
      }
      dq = 1.0 * P.dy2 * minmod3(s0, v1, s2, ys0, ys1, ys2);
      S[i] = v1 - dq;
      Nn[i] = v1 + dq;
      if (k == 0) {
        qp2_uv(S, u2, w2);
        S[i] = ii ? w2 : u2;
      } else if (k == ny - 1) {
        qp2_uv(Nn, u2, w2);
        Nn[i] = ii ? w2 : u2;
      }
    }
    const size_t fw = (size_t)j + (size_t)(nx + 1) * k, fe = fw + 1;
    const size_t fs = (size_t)j + (size_t)nx * k, fn = fs + nx;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      B.qpR[i * NFX + fw] = W[i];
      B.qpL[i * NFX + fe] = E[i];
      B.qpT[i * NFY + fs] = S[i];
      B.qpB[i * NFY + fn] = Nn[i];
      if (j == 0) B.qpL[i * NFX + fw] = W[i];
      if (j == nx - 1) B.qpR[i * NFX + fe] = E[i];
      if (k == 0) B.qpB[i * NFY + fs] = S[i];
      if (k == ny - 1) B.qpT[i * NFY + fn] = Nn[i];
    }
  }
}

// eval_speeds + eval_flux_KT of one face (src/solver_2d.f90:2834-3049,
// :3937-4014): x faces first (f < NFX), then y faces
template <int NS, int NX>
__global__ __launch_bounds__(PTHR) void k_pp_faces(Params P, Grid G, PoreBufs B) {
  constexpr int NV = 4 + NS + nx_count(NX), NP = NV + 2;
  if (step_done(G)) return;
  const int nx = G.nx, ny = G.ny;
  const int NFX = (nx + 1) * ny, NFY = nx * (ny + 1);
  for (int e = blockIdx.x * PTHR + threadIdx.x; e < NFX + NFY; e += gridDim.x * PTHR) {
    const bool xd = e < NFX;
    const int f = xd ? e : e - NFX, nf = xd ? NFX : NFY, dir = xd ? 1 : 2;
    const double* sa = xd ? B.qpL : B.qpB;
    const double* sb = xd ? B.qpR : B.qpT;
    double pa[NP], pb[NP], qa[NV], qb[NV], fa[NV], fb[NV];
#pragma unroll
    for (int i = 0; i < NP; i++) {
      pa[i] = sa[(size_t)i * nf + f];
      pb[i] = sb[(size_t)i * nf + f];
    }
    qp_to_qc<NS, double, NX>(P, pa, qa);
    qp_to_qc<NS, double, NX>(P, pb, qb);
    double amin, amax, bmin, bmax;
    local_speeds<NS, double, NX>(P, pa, dir, amin, amax);
    local_speeds<NS, double, NX>(P, pb, dir, bmin, bmax);
    const double an = fmin_f(fmin_f(amin, bmin), 0.0), ap = fmax_f(fmax_f(amax, bmax), 0.0);
    phys_flux<NS, double, NX>(P, qa, pa, dir, fa);
    phys_flux<NS, double, NX>(P, qb, pb, dir, fb);
    double H[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
      if (an == ap) {
        H[i] = 0.0;
      } else {
        const double avg = (ap * fa[i] - an * fb[i]) / (ap - an);
        H[i] = avg + (ap * an) / (ap - an) * (qb[i] - qa[i]);
      }
    }
    double Ssum = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) Ssum += H[4 + i];
    if (xd) {
      if ((Ssum / H[0]) >= 1.0) {
        double S2 = 0.0;
#pragma unroll
        for (int i = 0; i < NS; i++) S2 += H[4 + i] / H[0];
#pragma unroll
        for (int i = 0; i < NS; i++) H[4 + i] = H[4 + i] / S2;
      }
      if ((pa[1] == 0.0) && (pb[1] == 0.0)) {
        H[0] = 0.0;
#pragma unroll
        for (int i = 3; i < NV; i++) H[i] = 0.0;
      }
    } else {
      if ((Ssum / H[0]) > 1.0) {
        const double S2 = Ssum / H[0];
#pragma unroll
        for (int i = 0; i < NS; i++) H[4 + i] = H[4 + i] / S2;
      }
      if ((qa[2] == 0.0) && (qb[2] == 0.0)) {
        H[0] = 0.0;
#pragma unroll
        for (int i = 3; i < NV; i++) H[i] = 0.0;
      }
    }
    double* Hd = xd ? B.Hx : B.Hy;
#pragma unroll
    for (int i = 0; i < NV; i++) Hd[(size_t)i * nf + f] = H[i];
    (xd ? B.amx : B.amy)[f] = fmax_f(ap, -an);
  }
}

// active cells: the stage-1 right-hand side divF - expl (eval_hyperbolic_terms
// :2616-2693, eval_expl_terms) into G.R and the CFL candidate of the cell
// (timestep, :873-1006) as one atomicMin on its bit pattern
template <int NS, int NX>
__global__ __launch_bounds__(PTHR) void k_pp_cells(Params P, Grid G, PoreBufs B) {
  constexpr int NV = 4 + NS + nx_count(NX), NP = NV + 2;
  if (step_done(G)) return;
  const int nx = G.nx, ny = G.ny, N = G.N;
  const size_t NFX = (size_t)(nx + 1) * ny, NFY = (size_t)nx * (ny + 1);
  for (int c = blockIdx.x * PTHR + threadIdx.x; c < N; c += gridDim.x * PTHR) {
    if (!(G.mask[c] & 1u)) continue;
    const int j = c % nx, k = c / nx;
    const size_t fw = (size_t)j + (size_t)(nx + 1) * k, fe = fw + 1;
    const size_t fs = (size_t)j + (size_t)nx * k, fn = fs + nx;
    double qp[NP], ex[NV];
#pragma unroll
    for (int i = 0; i < NP; i++) qp[i] = B.qp[(size_t)i * N + c];
    expl_terms<NS, double, NX>(P, G.bx[c], G.by[c], qp, ex, G.st->t, 0.0, 0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double d = 0.0;
      d = d + (B.Hx[i * NFX + fe] - B.Hx[i * NFX + fw]) * P.one_by_dx;
      d = d + (B.Hy[i * NFY + fn] - B.Hy[i * NFY + fs]) * P.one_by_dy;
      G.R[(size_t)i * N + c] = d - ex[i];
    }
    // a cell whose faces are all still contributes nothing (the reference
    // takes its running minimum there)
    const double inf = __longlong_as_double(0x7ff0000000000000LL);
    double ma = fmax_f(B.amx[fw], B.amx[fe]);
    const double dtx = (ma > 0.0) ? P.cfl * P.dx / ma : inf;
    ma = fmax_f(B.amy[fs], B.amy[fn]);
    const double dty = (ma > 0.0) ? P.cfl * P.dy / ma : inf;
    const double dtc = fmin_f(dtx, dty);
    if (dtc < inf) atomicMin(&G.st->dtmin_bits, (unsigned long long)__double_as_longlong(dtc));
  }
}

// ---------------------------------------------------------------------------
// stage 2 with the Newton system (src/solver_2d.f90:1103-1338,
// :1662-2367): hu, hv and the pore pressure are implicit (n_nh = 3, solved
// by DGESV), or hu and hv alone (n_nh = 2, Cramer's rule); the solid classes
// and the transported noise are explicit rows
// ---------------------------------------------------------------------------
template <int NS, int NX>
__device__ __forceinline__ bool implicit_var(int i) {
  return i == 1 || i == 2 || (has_pore(NX) && i == pore_at<NS, NX>());
}

// eval_implicit_terms at a real state: the complex code on (q, 0)
template <int NS, int NX>
__device__ void implicit_real(const Params& P, double xi, const double* q, double* nh) {
  constexpr int NV = 4 + NS + nx_count(NX);
  cx qc[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) qc[i] = C(q[i], 0.0);
  cx n1, n2, np;
  implicit_c<NS, NX>(P, qc, n1, n2, xi, &np);
#pragma unroll
  for (int i = 0; i < NV; i++) nh[i] = 0.0;
  nh[1] = n1.re;
  nh[2] = n2.re;
  if constexpr (has_pore(NX)) nh[pore_at<NS, NX>()] = np.re;
}

template <int NS, int NX>
__device__ void eval_f_pp(const Params& P, double xi, double dt, const double* qj, const double* q0,
                          const double* Rni, double* f, double& scal) {
  constexpr int NV = 4 + NS + nx_count(NX);
  double nh[NV];
  implicit_real<NS, NX>(P, xi, qj, nh);
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double Rj = Rni[i] - 1.0 * nh[i];
    f[i] = qj[i] - q0[i] + dt * Rj;
    f[i] = 1.0 * f[i];
  }
#pragma unroll
  for (int i = 0; i < NV; i++) s += f[i] * f[i];
  scal = 0.5 * s;
}

// eval_jacobian (complex step, :2302-2367): L column-major NV x NV
template <int NS, int NX>
__device__ void eval_jacobian_pp(const Params& P, double xi, double dt, const double* qrel,
                                 const double* qorg, double* L) {
  constexpr int NV = 4 + NS + nx_count(NX);
#pragma unroll
  for (int i = 0; i < NV * NV; i++) L[i] = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    L[i + NV * i] = 1.0 * qorg[i];
    if (implicit_var<NS, NX>(i)) {
      cx qc[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) qc[m] = cmul(C(qrel[m], (m == i) ? P.hcs : 0.0), C(qorg[m], 0.0));
      cx n1, n2, np;
      implicit_c<NS, NX>(P, qc, n1, n2, xi, &np);
#pragma unroll
      for (int r = 0; r < NV; r++) {
        const double im =
            (r == 1) ? n1.im
                     : ((r == 2) ? n2.im
                                 : ((has_pore(NX) && r == pore_at<NS, NX>()) ? np.im : 0.0));
        const double J = 1.0 * im * P.one_by_hcs;
        L[r + NV * i] = L[r + NV * i] - dt * 1.0 * J;
      }
    }
  }
}

// DGESV of the implicit block as the reference links it (SciPy's OpenBLAS):
// a left-looking (Crout) LU with partial pivoting -- column j takes the
// earlier row interchanges, a(i,j) -= dot(a(i,0:m), a(0:m,j)) accumulated by
// fused multiply-adds from 0 (m = i above the diagonal, j below), the pivot
// is the first largest |a(i,j)|, rows j and p swap in columns 0..j, and the
// column below the pivot is scaled by the pivot's reciprocal -- then the
// permuted forward and backward solves with fused multiply-adds.  A zero
// pivot leaves b unchanged, as DGESV does.  a: column-major 3 x 3.
__device__ void dgesv3(double* a, double* b) {
  constexpr int n = 3;
  int ipiv[n];
  int info = 0;
#pragma unroll
  for (int j = 0; j < n; j++) {
#pragma unroll
    for (int i = 0; i < j; i++) {
      const int p = ipiv[i];
      if (p != i) {
        const double t = a[i + n * j];
        a[i + n * j] = a[p + n * j];
        a[p + n * j] = t;
      }
    }
#pragma unroll
    for (int i = 1; i < j; i++) {
      double d = 0.0;
#pragma unroll
      for (int k = 0; k < i; k++) d = fma(a[i + n * k], a[k + n * j], d);
      a[i + n * j] -= d;
    }
#pragma unroll
    for (int i = j; i < n; i++) {
      double d = 0.0;
#pragma unroll
      for (int k = 0; k < j; k++) d = fma(a[i + n * k], a[k + n * j], d);
      a[i + n * j] -= d;
    }
    int p = j;
    double mx = fabs(a[j + n * j]);
#pragma unroll
    for (int i = j + 1; i < n; i++)
      if (fabs(a[i + n * j]) > mx) {
        mx = fabs(a[i + n * j]);
        p = i;
      }
    ipiv[j] = p;
    if (p != j) {
#pragma unroll
      for (int k = 0; k <= j; k++) {
        const double t = a[j + n * k];
        a[j + n * k] = a[p + n * k];
        a[p + n * k] = t;
      }
    }
    if (a[j + n * j] != 0.0) {
      const double r = 1.0 / a[j + n * j];
#pragma unroll
      for (int i = j + 1; i < n; i++) a[i + n * j] *= r;
    } else if (!info) {
      info = j + 1;
    }
  }
  if (info) return;
#pragma unroll
  for (int j = 0; j < n; j++) {
    const int p = ipiv[j];
    if (p != j) {
      const double t = b[j];
      b[j] = b[p];
      b[p] = t;
    }
  }
#pragma unroll
  for (int j = 0; j < n; j++)
#pragma unroll
    for (int i = j + 1; i < n; i++) b[i] = fma(-b[j], a[i + n * j], b[i]);
#pragma unroll
  for (int j = n - 1; j >= 0; j--) {
    b[j] /= a[j + n * j];
#pragma unroll
    for (int i = 0; i < j; i++) b[i] = fma(-b[j], a[i + n * j], b[i]);
  }
}

// lnsrch (:2021-2226), one thread per cell
template <int NS, int NX>
__device__ void lnsrch_pp(const Params& P, double xi, double dt, const double* xold,
                          const double* qorg, const double* q0, double fold, const double* g,
                          double* p, double* x, double& f_out, double* rt, double stpmax,
                          const double* Rni, unsigned int& err) {
  constexpr int NV = 4 + NS + nx_count(NX);
  const double TOLX = kEps;
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) sum += p[i] * p[i];
  const double pabs = sqrt(sum);
  if (pabs > stpmax) {
#pragma unroll
    for (int i = 0; i < NV; i++) p[i] = p[i] * stpmax / pabs;
  }
  double slope = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) slope += g[i] * p[i];
  double test = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double tmp = fabs(p[i]) / fmax_f(fabs(xold[i]), 1.0);
    if (i == 0 || tmp > test) test = tmp;
  }
  const double alamin = TOLX / test;
  if (alamin == 0.0) {
#pragma unroll
    for (int i = 0; i < NV; i++) x[i] = xold[i];
    return;  // rt and f_out still hold the residual at xold
  }
  double alam = 1.0, alam2 = 0.0, f2 = 0.0, tmplam = 0.0, f = 0.0;
  double qj[NV];
  for (int trial = 0;; trial++) {
#pragma unroll
    for (int i = 0; i < NV; i++) x[i] = xold[i] + alam * p[i];
#pragma unroll
    for (int i = 0; i < NV; i++) qj[i] = x[i] * qorg[i];
    eval_f_pp<NS, NX>(P, xi, dt, qj, q0, Rni, rt, f);
    if (f <= 0.9 * fold) {
      break;
    } else if (alam < alamin) {
#pragma unroll
      for (int i = 0; i < NV; i++) x[i] = xold[i];
      f = fold;
      break;
    } else {
      if (alam == 1.0) {
        tmplam = -(slope / (2.0 * (f - fold - slope)));
      } else {
        const double rhs1 = f - fold - alam * slope;
        const double rhs2 = f2 - fold - alam2 * slope;
        const double a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
        const double b = (-(alam2 * rhs1 / (alam * alam)) + alam * rhs2 / (alam2 * alam2)) /
                         (alam - alam2);
        if (a == 0.0) {
          tmplam = -(slope / (2.0 * b));
        } else {
          const double disc = b * b - 3.0 * a * slope;
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
    // the reference loop ends by construction (alam at least halves); a NaN
    // alamin would loop for ever: bounded, as an error
    if (trial >= 2200) {
      err |= IMEX_ERR_NAN;
      break;
    }
    alam2 = alam;
    f2 = f;
    alam = fmax_f(tmplam, 0.5 * alam);
  }
  f_out = f;
}

template <int NS, int NX>
__device__ void solve_rk_step_pp(const Params& P, double xi, double dt, double* qj, const double* q0,
                                 const double* Rni, unsigned int& err) {
  constexpr int NV = 4 + NS + nx_count(NX), NH = has_pore(NX) ? 3 : 2, NE = NV - NH;
  const double TOLF = 1.0e-10, STPMX = 100.0;
  double qorg[NV], qrel[NV], rt[NV], xold[NV], p[NV], grad[NV], L[NV * NV];
  double scal_f = 0.0, scal_f_old;
#pragma unroll
  for (int i = 0; i < NV; i++) grad[i] = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) qorg[i] = fmax_f(fabs(qj[i]), 1.0e-3);
#pragma unroll
  for (int i = 0; i < NV; i++) qrel[i] = qj[i] / qorg[i];
  bool fresh = false;
  for (int it = 1; it <= 100; it++) {
    if (!fresh) eval_f_pp<NS, NX>(P, xi, dt, qj, q0, Rni, rt, scal_f);
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mx = fmax_f(mx, fabs(rt[i]));
    if (mx < TOLF) break;
    eval_jacobian_pp<NS, NX>(P, xi, dt, qrel, qorg, L);
    // explicit rows by their diagonal, the implicit block by DGESV or
    // Cramer's rule (:1851-1926)
    double d1[NE], d2[NH];
    int imp[NH];
    int ne = 0, nn = 0;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      if (implicit_var<NS, NX>(r)) {
        imp[nn] = r;
        d2[nn] = rt[r];
        nn++;
      } else {
        d1[ne] = rt[r] / L[r + NV * r];
        ne++;
      }
    }
    {
      // MATMUL(d1, reshape(pack(L, mask21), [NE, NH])): the implicit rows of
      // the explicit columns, in pack order
      double pk[NE * NH];
      int m = 0;
#pragma unroll
      for (int col = 0; col < NV; col++) {
        if (implicit_var<NS, NX>(col)) continue;
#pragma unroll
        for (int b = 0; b < NH; b++) pk[m++] = L[imp[b] + NV * col];
      }
#pragma unroll
      for (int b = 0; b < NH; b++) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NE; a++) s += d1[a] * pk[a + NE * b];
        d2[b] = d2[b] - s;
      }
    }
    double a22[NH * NH];
#pragma unroll
    for (int cc = 0; cc < NH; cc++)
#pragma unroll
      for (int r = 0; r < NH; r++) a22[r + NH * cc] = L[imp[r] + NV * imp[cc]];
    if constexpr (NH == 3) {
      dgesv3(a22, d2);
    } else {
      // :1891-1901
      const double inv_det = 1.0 / (a22[0] * a22[3] - a22[1] * a22[2]);
      const double s1 = (d2[0] * a22[3] - d2[1] * a22[2]) * inv_det;
      const double s2 = (a22[0] * d2[1] - a22[1] * d2[0]) * inv_det;
      d2[0] = s1;
      d2[1] = s2;
    }
    ne = 0;
    nn = 0;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      if (implicit_var<NS, NX>(r)) {
        p[r] = -d2[nn] + 0.0;
        nn++;
      } else {
        p[r] = 0.0 + -d1[ne];
        ne++;
      }
    }
#pragma unroll
    for (int i = 0; i < NV; i++) xold[i] = qrel[i];
    scal_f_old = scal_f;
    if (it > 1) {
      double ss = 0.0;
#pragma unroll
      for (int i = 0; i < NV; i++) ss += qrel[i] * qrel[i];
      const double stpmax = STPMX * fmax_f(sqrt(ss), (double)NV);
#pragma unroll
      for (int i = 0; i < NV; i++) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < NV; r++) s += rt[r] * L[r + NV * i];
        grad[i] = s;
      }
      lnsrch_pp<NS, NX>(P, xi, dt, xold, qorg, q0, scal_f_old, grad, p, qrel, scal_f, rt, stpmax, Rni,
                    err);
      fresh = true;
    } else {
#pragma unroll
      for (int i = 0; i < NV; i++) qrel[i] = xold[i] + p[i];
#pragma unroll
      for (int i = 0; i < NV; i++) qj[i] = qrel[i] * qorg[i];
      eval_f_pp<NS, NX>(P, xi, dt, qj, q0, Rni, rt, scal_f);
      fresh = true;
    }
#pragma unroll
    for (int i = 0; i < NV; i++) qj[i] = qrel[i] * qorg[i];
    mx = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mx = fmax_f(mx, fabs(rt[i]));
    if (mx < TOLF) break;
    double test = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      const double tmp = fabs(qrel[i] - xold[i]) / fmax_f(fabs(qrel[i]), 1.0);
      if (i == 0 || tmp > test) test = tmp;
    }
    if (test < kEps) break;
  }
}

// stage 2 and assemble_sol for one cell; q holds q0 on entry
template <int NS, int NX>
__device__ void stage2_pp(const Params& P, double xi, double dt, double* q, const double* X,
                          unsigned int& err) {
  constexpr int NV = 4 + NS + nx_count(NX), NP = NV + 2;
  double q0[NV], qfv[NV], qp[NP], q_si[NV], qg[NV], si[NV], Rj[NV], pd;
#pragma unroll
  for (int i = 0; i < NV; i++) q0[i] = q[i];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double A = 0.0;
    A = A + X[i] * 1.0;
    A = A + (0.0 - 0.0) * 0.0;
    double Bv = 0.0;
    Bv = Bv + (0.0 + 0.0) * 0.0;
    Bv = Bv + (0.0 + 0.0) * 0.0;
    qfv[i] = q0[i] - dt * (A - Bv);
  }
  qc_to_qp<NS, double, NX>(P, qfv, qp, pd);
  if (qfv[0] > 0.0) {
    semi_impl<NS, NX>(P, qp, si);
#pragma unroll
    for (int i = 0; i < NV; i++) q_si[i] = qfv[i] + dt * 1.0 * si[i];
    if ((qfv[1] * qfv[1] + qfv[2] * qfv[2]) == 0.0) {
      q_si[1] = 0.0;
      q_si[2] = 0.0;
    } else if ((q_si[1] * qfv[1] < 0.0) || (q_si[2] * qfv[2] < 0.0)) {
      q_si[1] = 0.0;
      q_si[2] = 0.0;
    } else {
      const double m = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
      const double d = sqrt(qfv[1] * qfv[1] + qfv[2] * qfv[2]);
      const double a1 = m * qfv[1] / d, a2 = m * qfv[2] / d;
      q_si[1] = a1;
      q_si[2] = a2;
    }
#pragma unroll
    for (int i = 0; i < NV; i++) si[i] = (q_si[i] - qfv[i]) / (dt * 1.0);
#pragma unroll
    for (int i = 0; i < NV; i++) qg[i] = q_si[i];
#pragma unroll
    for (int i = 0; i < NV; i++) {
      const double A = 0.0 + X[i] * 1.0;
      const double Bv = 0.0 + (0.0 + 0.0) * 0.0;
      Rj[i] = (A - Bv) - 1.0 * si[i];
    }
    solve_rk_step_pp<NS, NX>(P, xi, dt, qg, q0, Rj, err);
    if (q_si[1] * q_si[1] + q_si[2] * q_si[2] == 0.0) {
      qg[1] = 0.0;
      qg[2] = 0.0;
    } else if ((qg[1] * q_si[1] <= 0.0) && (qg[2] * q_si[2] <= 0.0)) {
      qg[1] = 0.0;
      qg[2] = 0.0;
    } else {
      const double m = sqrt(qg[1] * qg[1] + qg[2] * qg[2]);
      const double d = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
      const double a1 = m * q_si[1] / d, a2 = m * q_si[2] / d;
      qg[1] = a1;
      qg[2] = a2;
    }
  } else {
#pragma unroll
    for (int i = 0; i < NV; i++) qg[i] = qfv[i];
  }
  // assemble_sol (:1342-1630)
#pragma unroll
  for (int i = 0; i < NV; i++) q[i] = qg[i];
#pragma unroll
  for (int i = 0; i < NV; i++)
    if (isnan(q[i])) err |= IMEX_ERR_NAN;
  if (q[0] < kEps) {
    if (q[0] > -1.0e-7) {
#pragma unroll
      for (int i = 0; i < NV; i++) q[i] = 0.0;
    } else {
      err |= IMEX_ERR_NEG_H;
    }
  }
  bool anyneg = false, anybad = false;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    if (q[4 + i] < 0.0) anyneg = true;
    if (q[4 + i] <= -1.0e-7) anybad = true;
  }
  if (anyneg) {
    if (anybad) {
      err |= IMEX_ERR_NEG_SOLID;
    } else {
#pragma unroll
      for (int i = 0; i < NS; i++)
        if (q[4 + i] > -1.0e-7) q[4 + i] = 0.0;
    }
  }
  qc_to_qp<NS, double, NX>(P, q, qp, pd);
  if (qp[3] < 273.0) err |= IMEX_ERR_COLD;
  double S = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) S += q[4 + i];
  if (S > q[0]) {
    double S2 = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) S2 += q[4 + i];
    const bool ok = ((S2 - q[0]) / q[0] < 1.0e-10) || (q[0] < kEps);
    if (!ok) err |= IMEX_WARN_RESCALE;
    double S3 = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) S3 += q[4 + i];
#pragma unroll
    for (int i = 0; i < NS; i++) q[4 + i] = q[4 + i] / S3 * q[0];
  }
}

// update_erosion_deposition_cell (:2382-2597) with the mass exchange of the
// pore-pressure runs (gas loss; no settling, entrainment or erosion here)
template <int NS, int NX>
__device__ void mass_exchange_pp(const Params& P, double dt, double* q, double* depo,
                                 unsigned int& err) {
  constexpr int NV = 4 + NS + nx_count(NX), NP = NV + 2;
  double qp[NP], pd, dep[NS], eq[NV], ers[NS], erod[NS], topo;
#pragma unroll
  for (int i = 0; i < NS; i++) erod[i] = 0.0;
  if (q[0] > 0.0) {
    qc_to_qp<NS, double, NX>(P, q, qp, pd);
  } else {
#pragma unroll
    for (int i = 0; i < NP; i++) qp[i] = 0.0;
    qp[3] = P.T_amb;
  }
  if (mass_exchange<NS, 1, NX>(P, qp, erod, dt, dep, eq, ers, topo, nullptr))
    err |= IMEX_ERR_NEG_DEP;
#pragma unroll
  for (int i = 0; i < NV; i++) q[i] = q[i] + dt * eq[i];
#pragma unroll
  for (int i = 0; i < NS; i++) q[4 + i] = fmax_f(0.0, q[4 + i]);
#pragma unroll
  for (int i = 0; i < NS; i++) depo[i] = depo[i] + dt * dep[i];
#pragma unroll
  for (int i = 0; i < NS; i++)
    if (q[4 + i] < 0.0) err |= IMEX_ERR_NEG_SOLID;
  if (q[0] <= 0.0) {
    if (q[0] > -1.0e-10) {
#pragma unroll
      for (int i = 0; i < NV; i++) q[i] = 0.0;
    } else {
      err |= IMEX_ERR_NEG_H;
    }
  }
  double S = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) S += q[4 + i];
  if (S > q[0]) {
    if (q[0] < 1.0e-10) {
      double S2 = 0.0;
#pragma unroll
      for (int i = 0; i < NS; i++) S2 += q[4 + i];
#pragma unroll
      for (int i = 0; i < NS; i++) q[4 + i] = q[4 + i] / S2 * q[0];
    } else {
      err |= IMEX_ERR_SOLID_GT_TOT;
    }
  }
  double rg;
  if (q[0] > 0.0) {
    qc_to_qp<NS, double, NX>(P, q, qp, pd);
    rg = mixt_var<NS, double, NX>(P, qp, false).red_grav;
  } else {
    rg = 0.0;
  }
  if (rg <= 0.0) {
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = 0.0;
  }
}

// driver epilogue (src/IMEX_SfloW2D.f90:449-505): running maxima and the
// vulnerability bits of one cell
template <int NS, int NX>
__device__ void epilogue_pp(const Params& P, const double* q, double& hmax, double& pdynmax,
                            double& mvel, unsigned long long& vuln) {
  constexpr int NV = 4 + NS + nx_count(NX), NP = NV + 2;
  if (!(q[0] > 0.0)) return;
  double qp[NP], pd;
  qc_to_qp<NS, double, NX>(P, q, qp, pd);
  hmax = fmax_f(hmax, qp[0]);
  const double ru = qp[NV], rv = qp[NV + 1];
  const double mv = sqrt(ru * ru + rv * rv);
  if (qp[0] > 0.001) {
    pdynmax = fmax_f(pdynmax, pd);
    mvel = fmax_f(mvel, mv);
  }
  int it = 0;
  unsigned long long bits = 0ULL;
  for (int a = 0; a < P.n_thk; a++) {
    const bool thk = (qp[0] >= P.thk[a]);
    for (int b = 0; b < P.n_pdyn; b++) {
      const bool pdn = (pd >= P.pdyn[b]);
      if (thk && pdn) bits |= (1ULL << it);
      it++;
    }
  }
  vuln = vuln | bits;
}

template <int NS, int NX>
__global__ __launch_bounds__(PTHR) void k_pp_update(Params P, Grid G, int step_idx) {
  constexpr int NV = 4 + NS + nx_count(NX);
  DevState* st = G.st;
  if (st->done) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && G.dts) G.dts[step_idx] = 0.0;
    return;
  }
  const double dt = driver_dt(P, st);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->stop_next = (st->t + dt < st->t_end) ? 0u : 1u;
    st->dt = dt;
    st->dt_cfl = fmin_f(P.max_dt, __longlong_as_double((long long)st->dtmin_bits));
    if (G.dts) G.dts[step_idx] = dt;
    st->pending = 1u;  // committed by the next launch on the stream
  }
  const int N = G.N;
  unsigned int err = 0;
  int dried = 0, ncell = 0;
  for (int c = blockIdx.x * PTHR + threadIdx.x; c < N; c += gridDim.x * PTHR) {
    if (!(G.mask[c] & 1u)) continue;
    ncell++;
    double q[NV], X[NV], depo[NS];
#pragma unroll
    for (int v = 0; v < NV; v++) q[v] = G.q[(size_t)v * N + c];
#pragma unroll
    for (int v = 0; v < NV; v++) X[v] = G.R[(size_t)v * N + c];
#pragma unroll
    for (int v = 0; v < NS; v++) depo[v] = G.dep[(size_t)v * N + c];
    const bool wet0 = q[0] > 0.0;
    stage2_pp<NS, NX>(P, cell_xi(P, G.noise, c), dt, q, X, err);
    // update_erosion_deposition_cell returns at once without erosion,
    // settling, entrainment and pore pressure (src/solver_2d.f90:2439-2441)
    if constexpr (has_pore(NX)) mass_exchange_pp<NS, NX>(P, dt, q, depo, err);
    if (wet0 && q[0] == 0.0) dried++;
    double hm = G.hmax[c], pm = G.pdyn[c], mv = G.mvel[c];
    unsigned long long vb = G.vuln ? G.vuln[c] : 0ULL;
    epilogue_pp<NS, NX>(P, q, hm, pm, mv, vb);
#pragma unroll
    for (int v = 0; v < NV; v++) G.q[(size_t)v * N + c] = q[v];
#pragma unroll
    for (int v = 0; v < NS; v++) G.dep[(size_t)v * N + c] = depo[v];
    G.hmax[c] = hm;
    G.pdyn[c] = pm;
    G.mvel[c] = mv;
    if (G.vuln) G.vuln[c] = vb;
  }
  if (err) atomicOr(&st->err, err);
  if (dried) atomicAdd(&st->dried, (unsigned long long)dried);
  if (ncell) atomicAdd(&st->work, (unsigned long long)ncell);
}

template <int NS, int NX>
int launch_step(imex_handle* h, const PoreBufs& B, int step_idx) {
  const Grid& G = h->G;
  const Params& P = h->P;
  const int nx = G.nx, ny = G.ny;
  const int nfaces = (nx + 1) * ny + nx * (ny + 1);
  const int gc = std::max(1, std::min((G.N + PTHR - 1) / PTHR, 4 * h->n_cu));
  const int gf = std::max(1, std::min((nfaces + PTHR - 1) / PTHR, 4 * h->n_cu));
  hipLaunchKernelGGL((k_pp_prim<NS, NX>), dim3(gc), dim3(PTHR), 0, h->stream, P, G, B);
  hipLaunchKernelGGL((k_pp_recon<NS, NX>), dim3(gc), dim3(PTHR), 0, h->stream, P, G, B);
  hipLaunchKernelGGL((k_pp_faces<NS, NX>), dim3(gf), dim3(PTHR), 0, h->stream, P, G, B);
  hipLaunchKernelGGL((k_pp_cells<NS, NX>), dim3(gc), dim3(PTHR), 0, h->stream, P, G, B);
  hipLaunchKernelGGL((k_pp_update<NS, NX>), dim3(gc), dim3(PTHR), 0, h->stream, P, G, step_idx);
  HCHK(h, hipGetLastError());
  return 0;
}

}  // namespace

int imex_pore_alloc(imex_handle* h) {
  const size_t N = (size_t)h->G.N, nx = (size_t)h->G.nx, ny = (size_t)h->G.ny;
  const size_t NFX = (nx + 1) * ny, NFY = nx * (ny + 1);
  const size_t nv = (size_t)h->P.nv, np = nv + 2;
  int rc;
  if ((rc = dalloc(h, &h->pp_qp, np * N))) return rc;
  if ((rc = dalloc(h, &h->pp_qpL, np * NFX))) return rc;
  if ((rc = dalloc(h, &h->pp_qpR, np * NFX))) return rc;
  if ((rc = dalloc(h, &h->pp_qpB, np * NFY))) return rc;
  if ((rc = dalloc(h, &h->pp_qpT, np * NFY))) return rc;
  if ((rc = dalloc(h, &h->pp_Hx, nv * NFX))) return rc;
  if ((rc = dalloc(h, &h->pp_Hy, nv * NFY))) return rc;
  if ((rc = dalloc(h, &h->pp_amx, NFX))) return rc;
  if ((rc = dalloc(h, &h->pp_amy, NFY))) return rc;
  return 0;
}

// one device-driven step of a run on the general path (pore pressure, the
// transported noise or both) after the active-set pass
int imex_pore_step(imex_handle* h, int step_idx) {
  const PoreBufs B{h->pp_qp, h->pp_qpL, h->pp_qpR, h->pp_qpB, h->pp_qpT,
                   h->pp_Hx, h->pp_Hy, h->pp_amx, h->pp_amy};
  const int x = (h->P.pore ? 1 : 0) | (h->P.stoch_tr ? 2 : 0);
  switch (h->P.ns * 4 + x) {
    case 5: return launch_step<1, 1>(h, B, step_idx);
    case 6: return launch_step<1, 2>(h, B, step_idx);
    case 7: return launch_step<1, 3>(h, B, step_idx);
    case 9: return launch_step<2, 1>(h, B, step_idx);
    case 10: return launch_step<2, 2>(h, B, step_idx);
    case 11: return launch_step<2, 3>(h, B, step_idx);
    case 13: return launch_step<3, 1>(h, B, step_idx);
    case 14: return launch_step<3, 2>(h, B, step_idx);
    default:
      h->err = "general path: no kernels for this class count and extra variables";
      return -2;
  }
}
