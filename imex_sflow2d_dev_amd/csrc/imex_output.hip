// Device-side outputs of the driver: the grid reductions of output_runout
// (src/inpout_2d.f90:7170-7427) and the probe interpolation of output_probes
// (src/inpout_2d.f90:6926-7155 with interp_2d_scalarB, src/geometry_2d.f90:
// 1535-1599).  The reference computes both on the host over module arrays
// every output time; here one pass over the device state replaces the
// whole-array expressions, and the caller formats the files.
//
// Both read the primitive variables the driver keeps in its module array qp:
// qc_to_qp(q) where q(1) > 0, else zero with T = T_ambient
// (src/IMEX_SfloW2D.f90:318-360, :456-503).  Every cell whose state changed
// was in the active set of that step, so recomputing qp from q gives the
// module array's values.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <vector>

#include "imex_handle.h"

namespace {

constexpr int RB = 256;  // threads per reduction block

// One block's (or the whole grid's) share of the runout reductions.
struct RunoutPart {
  double maxh;
  double s[3], e[3];  // double-double sums of q1, X*q1, Y*q1
  unsigned long long nh, nold, nnew;
  double d, dxm, dym, dax, day, bm;  // maxima (dist, dist_x, dist_y), dist at the
                                     // dist_x / dist_y maxima, B of the wet cells
  long long id, idx, idy, ib;        // first linear index of each maximum
};

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& err) {
  s = a + b;
  const double bp = s - a;
  err = (a - (s - bp)) + (b - bp);
}

__device__ __forceinline__ void dd_add(double& s, double& e, double x) {
  double t, r;
  two_sum(s, x, t, r);
  s = t;
  e += r;
}

// MAXLOC: the larger value, and on a tie the first index in array order
// (x index fastest, as Fortran's column-major order of (j, k)).
__device__ __forceinline__ void amax(double& v, long long& i, double& aux, double v2, long long i2,
                                     double aux2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
    aux = aux2;
  }
}

__device__ __forceinline__ void part_init(RunoutPart& p) {
  p.maxh = -INFINITY;
  for (int m = 0; m < 3; m++) p.s[m] = p.e[m] = 0.0;
  p.nh = p.nold = p.nnew = 0;
  p.d = p.dxm = p.dym = p.bm = -INFINITY;
  p.dax = p.day = 0.0;
  p.id = p.idx = p.idy = p.ib = LLONG_MAX;
}

__device__ __forceinline__ void part_merge(RunoutPart& a, const RunoutPart& b) {
  a.maxh = (b.maxh > a.maxh) ? b.maxh : a.maxh;
  for (int m = 0; m < 3; m++) {
    double t, r;
    two_sum(a.s[m], b.s[m], t, r);
    a.s[m] = t;
    a.e[m] = (a.e[m] + b.e[m]) + r;
  }
  a.nh += b.nh;
  a.nold += b.nold;
  a.nnew += b.nnew;
  double unused = 0.0;
  amax(a.d, a.id, unused, b.d, b.id, 0.0);
  amax(a.dxm, a.idx, a.dax, b.dxm, b.idx, b.dax);
  amax(a.dym, a.idy, a.day, b.dym, b.idy, b.day);
  amax(a.bm, a.ib, unused, b.bm, b.ib, 0.0);
}

__device__ void block_reduce(RunoutPart& mine, RunoutPart* out) {
  __shared__ RunoutPart sh[RB];
  const int tid = threadIdx.x;
  sh[tid] = mine;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (tid < s) part_merge(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  if (tid == 0) *out = sh[0];
}

// The primitive thickness and SUM(qp(5:4+n_solid)) of one cell, as the
// driver's module array qp holds them.
template <int NS, typename S>
__device__ __forceinline__ void cell_h_alphas(const Params& P, const Grid& G, int c, double q1,
                                              double& h, double& alphas) {
  constexpr int NV = 4 + NS;
  h = 0.0;
  alphas = 0.0;
  if (q1 > 0.0) {
    double q[NV], qp[NV + 2], pd;
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = (double)state_q<S>(G)[(size_t)i * G.N + c];
    qc_to_qp<NS>(P, q, qp, pd);
    h = qp[0];
#pragma unroll
    for (int i = 0; i < NS; i++) alphas += qp[4 + i];
  }
}

// One pass over the grid: hpos = (qp(1) > 1e-5) .AND. (alphas_tot >=
// alphas_threshold) (src/inpout_2d.f90:7283-7284, :7344-7345), its counts
// against the mask of the previous call, the mass-centre sums (:7279-7280),
// MAXVAL(qp(1)) (:7264), and MAXLOC of dist, dist_x and dist_y
// (:7349-7365); with `flags & 2` the MAXLOC of B_cent over the wet cells
// (:7293-7294).  Cells outside hpos hold dist = 0, as the reference's
// zero-initialised arrays.
template <int NS, typename S>
__global__ __launch_bounds__(RB) void k_runout(Params P, Grid G, const double* bcent,
                                               unsigned char* hpos_old, double x0, double y0,
                                               double thr, unsigned flags, RunoutPart* part) {
  RunoutPart p;
  part_init(p);
  for (int c = blockIdx.x * RB + threadIdx.x; c < G.N; c += gridDim.x * RB) {
    const int j = c % G.nx, k = c / G.nx;
    const double q1 = (double)state_q<S>(G)[c];
    double h, alphas;
    cell_h_alphas<NS, S>(P, G, c, q1, h, alphas);
    const bool hpos = (h > 1.0e-5) && (alphas >= thr);
    p.maxh = (h > p.maxh) ? h : p.maxh;
    const double X = G.xc[j], Y = G.yc[k];
    dd_add(p.s[0], p.e[0], q1);
    dd_add(p.s[1], p.e[1], X * q1);
    dd_add(p.s[2], p.e[2], Y * q1);
    const bool old = hpos_old[c] != 0;
    p.nh += hpos;
    p.nold += old;
    p.nnew += (hpos && !old);
    if (flags & IMEX_RUNOUT_STORE) hpos_old[c] = hpos ? 1 : 0;
    double dist = 0.0, dx = 0.0, dy = 0.0;
    if (hpos) {
      const double ax = X - x0, ay = Y - y0;
      dist = sqrt(ax * ax + ay * ay);
      dx = sqrt(ax * ax);
      dy = sqrt(ay * ay);
    }
    double unused = 0.0;
    amax(p.d, p.id, unused, dist, c, 0.0);
    amax(p.dxm, p.idx, p.dax, dx, c, dist);
    amax(p.dym, p.idy, p.day, dy, c, dist);
    if (flags & IMEX_RUNOUT_B) amax(p.bm, p.ib, unused, (h > 1.0e-5) ? bcent[c] : 0.0, c, 0.0);
  }
  block_reduce(p, part + blockIdx.x);
}

// The partials in a fixed order (thread t: partials t, t + RB, ...), then
// the block tree: the result does not depend on launch timing.
__global__ __launch_bounds__(RB) void k_runout_final(RunoutPart* part, int n) {
  RunoutPart p;
  part_init(p);
  for (int b = threadIdx.x; b < n; b += RB) part_merge(p, part[b]);
  block_reduce(p, part + n);
}

// interp_2d_scalarB's cell and weights for one point: 1-based (ix, iy) of
// the lower-left cell, the weights of its column and row, and whether the
// point is outside (f2 = 0).  The reference clamps ix to >= 0 and then reads
// x_comp(0), outside the array, for a point left of the first cell centre;
// here such a point is outside.
struct ProbeW {
  int ix, iy;
  double ax, ay;
  bool outside;
};

__device__ ProbeW probe_weights(const Grid& G, double x2, double y2) {
  ProbeW w;
  w.outside = false;
  if (G.nx > 1) {
    double f = floor((x2 - G.xc[0]) / (G.xc[1] - G.xc[0])) + 1.0;
    f = fmax(0.0, fmin(f, (double)(G.nx - 1)));
    w.ix = (int)f;
    if (w.ix == 0) {
      w.outside = true;
      w.ax = 0.0;
    } else {
      w.ax = (G.xc[w.ix] - x2) / (G.xc[w.ix] - G.xc[w.ix - 1]);
    }
  } else {
    w.ix = 1;
    w.ax = 0.0;
  }
  if (G.ny > 1) {
    double f = floor((y2 - G.yc[0]) / (G.yc[1] - G.yc[0])) + 1.0;
    f = fmax(1.0, fmin(f, (double)(G.ny - 1)));
    w.iy = (int)f;
    w.ay = (G.yc[w.iy] - y2) / (G.yc[w.iy] - G.yc[w.iy - 1]);
  } else {
    w.iy = 1;
    w.ay = 0.0;
  }
  if (w.ax < 0.0 || w.ax > 1.0 || w.ay < 0.0 || w.ay > 1.0) w.outside = true;
  return w;
}

// f2 of interp_2d_scalarB from the four corner values f(ix,iy), f(ix,iy+1),
// f(ix+1,iy), f(ix+1,iy+1), in the reference's order of operations.
__device__ __forceinline__ double probe_interp(const Grid& G, const ProbeW& w, const double* f) {
  if (w.outside) return 0.0;
  if (G.nx == 1) return w.ay * f[0] + (1.0 - w.ay) * f[1];
  if (G.ny == 1) return w.ax * f[0] + (1.0 - w.ax) * f[2];
  return w.ax * (w.ay * f[0] + (1.0 - w.ay) * f[1]) +
         (1.0 - w.ax) * (w.ay * f[2] + (1.0 - w.ay) * f[3]);
}

// output_probes for one probe per thread: the values after the time column
// (src/inpout_2d.f90:7040-7147) -- h, rho_m, T, u, v, alpha_s / h per class,
// the deposit per class when settling is on, p_dyn.
template <int NS, typename S>
__global__ void k_probes(Params P, Grid G, int n, const double* xy, double* out) {
  constexpr int NV = 4 + NS;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int nout = 6 + NS + (P.settling ? NS : 0);
  double* o = out + (size_t)p * nout;
  const ProbeW w = probe_weights(G, xy[2 * p], xy[2 * p + 1]);
  // corner cells (0-based): (ix, iy), (ix, iy+1), (ix+1, iy), (ix+1, iy+1) in
  // 1-based terms; a corner the branch for a one-cell-wide grid does not
  // read is replaced by the first one
  int cc[4] = {0, 0, 0, 0};
  if (!w.outside) {
    const int j = w.ix - 1, k = w.iy - 1;
    const int j1 = (G.nx > 1) ? j + 1 : j, k1 = (G.ny > 1) ? k + 1 : k;
    cc[0] = j + G.nx * k;
    cc[1] = j + G.nx * k1;
    cc[2] = j1 + G.nx * k;
    cc[3] = j1 + G.nx * k1;
  }
  double qp[4][NV + 2], q1[4];
  for (int m = 0; m < 4; m++) {
    double q[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = (double)state_q<S>(G)[(size_t)i * G.N + cc[m]];
    q1[m] = q[0];
    if (q[0] > 0.0) {
      double pd;
      qc_to_qp<NS>(P, q, qp[m], pd);
    } else {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) qp[m][i] = 0.0;
      qp[m][3] = P.T_amb;
    }
  }
  double f[4];
  auto field = [&](int i) {
    for (int m = 0; m < 4; m++) f[m] = qp[m][i];
    return probe_interp(G, w, f);
  };
  const double h = field(0);
  o[0] = h;
  double pdyn = 0.0;
  if (h > 1.0e-5) {
    for (int m = 0; m < 4; m++) f[m] = q1[m];
    const double rhom = probe_interp(G, w, f) / h;
    const double T = field(3), u = field(NV), v = field(NV + 1);
    o[1] = rhom;
    o[2] = T;
    o[3] = u;
    o[4] = v;
#pragma unroll
    for (int i = 0; i < NS; i++) o[5 + i] = field(4 + i) / h;
    pdyn = 0.5 * rhom * (u * u + v * v);
  } else {
    o[1] = 0.0;
    o[2] = P.T_amb;
    o[3] = 0.0;
    o[4] = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) o[5 + i] = 0.0;
  }
  if (P.settling) {
#pragma unroll
    for (int i = 0; i < NS; i++) {
      for (int m = 0; m < 4; m++) f[m] = G.dep[(size_t)i * G.N + cc[m]];
      o[5 + NS + i] = probe_interp(G, w, f);
    }
  }
  o[nout - 1] = pdyn;
}

// ---------------------------------------------------------------------------
// output_solution's records (src/inpout_2d.f90:5934-6423), one cell per
// thread, in the reference's order of operations: the .q_2d pass zeroes
// |q| < 1e-20 in the state (:6043-6054) and writes x, y, q; the _deposit_
// file gets the deposit as it is; the .p_2d pass (:6169-6402) computes
// qc_to_qp and mixt_var of the (zeroed) state, zeroes |deposit| < 1e-20 in
// place and writes its columns; with erosion the _erosion_ and _erodible_
// files get x, y and the layers (:6060-6069) and the .p_2d pass zeroes
// |erosion| < 1e-20 in place too.  Configured path: no added gases, liquid,
// stochastic or pore-pressure variables: shear velocity and Rouse numbers
// 0; without erosion the erosion and erodible columns are 0.  Voellmy
// rheology (model 1): inertial number 0, and mu_eff reads pore_pres(1) of a
// zero-size array in the reference; here the excess pore pressure is 0, so
// mu_eff = mu * MAX(0, 1 - 0 / (rho_m h g')) on wet cells.  mu(I) models 11
// and 12: mu_eff and the inertial number of :6277-6335.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int p2d_cols(int ns) { return 19 + 4 * ns; }

template <int NS, typename S>
__global__ void k_snapshot(Params P, Grid G, const double* bcent, unsigned what, double* sq,
                           double* sq2d, double* sdep2d, double* sp2d, double* sfld,
                           double* sero2d, double* serd2d) {
  constexpr int NV = 4 + NS;
  const int N = G.N;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    const int j = c % G.nx, k = c / G.nx;
    const double X = G.xc[j], Y = G.yc[k];
    double q[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = (double)state_q<S>(G)[(size_t)i * N + c];
    if (what & IMEX_SNAP_Q2D) {
      bool z = false, rest = false;
#pragma unroll
      for (int i = 0; i < NV; i++)
        if (fabs(q[i]) < 1.0e-20) {  // -0.0 becomes +0.0 too; stores only changed values
          if (q[i] != 0.0 || signbit(q[i])) {
            state_q<S>(G)[(size_t)i * N + c] = S(0.0);
            z = true;
          }
          q[i] = 0.0;
        }
      // a dry cell left with other non-zero components: the active-set
      // passes must then visit every tile with active cells (as after such
      // an upload); the flag stays on the device, in stream order
      if (z && !(q[0] > 0.0))
        for (int i = 1; i < NV; i++) rest = rest || q[i] != 0.0;
      if (rest) atomicOr(&G.st->hot, 1u);
      double* r = sq2d + (size_t)c * (2 + NV);
      r[0] = X;
      r[1] = Y;
#pragma unroll
      for (int i = 0; i < NV; i++) r[2 + i] = q[i];
    }
    double dep[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) dep[i] = G.dep[(size_t)i * N + c];
    if (what & IMEX_SNAP_DEP2D) {
      double* r = sdep2d + (size_t)c * (2 + NS);
      r[0] = X;
      r[1] = Y;
#pragma unroll
      for (int i = 0; i < NS; i++) r[2 + i] = dep[i];
    }
    const bool er = G.erosion != nullptr;
    double ero[NS], erd[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      ero[i] = er ? G.erosion[(size_t)i * N + c] : 0.0;
      erd[i] = er ? G.erodible[(size_t)i * N + c] : 0.0;
    }
    if (what & IMEX_SNAP_ERO2D) {
      double* r = sero2d + (size_t)c * (2 + NS);
      r[0] = X;
      r[1] = Y;
#pragma unroll
      for (int i = 0; i < NS; i++) r[2 + i] = ero[i];
    }
    if (what & IMEX_SNAP_ERD2D) {
      double* r = serd2d + (size_t)c * (2 + NS);
      r[0] = X;
      r[1] = Y;
#pragma unroll
      for (int i = 0; i < NS; i++) r[2 + i] = erd[i];
    }
    if (what & IMEX_SNAP_P2D) {
      double qp[NV + 2], pd;
      qc_to_qp<NS>(P, q, qp, pd);
      const MixtT<double> m = mixt_var<NS>(P, qp, false);
      double h = qp[0], u = qp[NV], v = qp[NV + 1], T = qp[3];
      double rho_m = m.rho_m, rg = m.red_grav;
      double al[NS];
      double mu_eff = 0.0, inertial = 0.0;
      if (h > 0.0) {
#pragma unroll
        for (int i = 0; i < NS; i++) al[i] = qp[4 + i] / h;
        if (P.rheology == 1) {
          const double excess = fmax_f(0.0, P.pres - P.pres);
          mu_eff = P.mu * fmax_f(0.0, (1.0 - excess / (rho_m * h * rg)));
        } else {
          // mu(I) models (:6277-6335): mu_s at rest, else mu(I) of the
          // horizontal speed; the thickness column shows the EPSILON floor
          const double mhv = sqrt(u * u + v * v);
          if (mhv < kEps) {
            mu_eff = P.mu_s;
          } else {
            if (h < kEps) h = kEps;
            const double vse = rho_m * rg * h;
            double diam = P.diam_s[0], rhop = P.rho_s[0];
            if (NS > 1) mean_particle<NS>(P, al, diam, rhop);
            const double shear = 5.0 / 2.0 * mhv / h;
            double ens = fmax_f(0.0, vse * sqrt(1.0));
            if (ens < kEps) ens = kEps;
            inertial = diam * shear / sqrt(ens / rhop);
            mu_eff = (P.mu_s * P.I_0 + P.mu_2 * inertial + P.muI_inf * (inertial * inertial)) /
                     (P.I_0 + inertial);
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < NS; i++) al[i] = 0.0;
      }
      auto tiny = [](double& x) {
        if (fabs(x) < 1.0e-20) x = 0.0;
      };
      tiny(h);
      tiny(u);
      tiny(v);
      double B = bcent ? bcent[c] : 0.0;
      tiny(B);
#pragma unroll
      for (int i = 0; i < NS; i++) {
        tiny(al[i]);
        if (fabs(dep[i]) < 1.0e-20) {
          if (dep[i] != 0.0 || signbit(dep[i])) G.dep[(size_t)i * N + c] = 0.0;
          dep[i] = 0.0;
        }
        if (fabs(ero[i]) < 1.0e-20) {
          if (er && (ero[i] != 0.0 || signbit(ero[i]))) G.erosion[(size_t)i * N + c] = 0.0;
          ero[i] = 0.0;
        }
      }
      double erd_sum = 0.0;
#pragma unroll
      for (int i = 0; i < NS; i++) erd_sum += erd[i];
      tiny(T);
      tiny(rho_m);
      tiny(rg);
      tiny(mu_eff);
      tiny(inertial);
      double* r = sp2d + (size_t)c * p2d_cols(NS);
      int o = 0;
      r[o++] = X;
      r[o++] = Y;
      r[o++] = h;
      r[o++] = u;
      r[o++] = v;
      r[o++] = B;
      r[o++] = h + B;
#pragma unroll
      for (int i = 0; i < NS; i++) r[o++] = al[i];
      r[o++] = T;
      r[o++] = rho_m;
      r[o++] = rg;
#pragma unroll
      for (int i = 0; i < NS; i++) r[o++] = dep[i];
#pragma unroll
      for (int i = 0; i < NS; i++) r[o++] = ero[i];
      r[o++] = erd_sum / (1.0 - P.erodible_porosity);
      r[o++] = 0.0;                                // alpha_l
      r[o++] = 0.0;                                // shear velocity
      r[o++] = m.Ri;
#pragma unroll
      for (int i = 0; i < NS; i++) r[o++] = 0.0;  // Rouse numbers
      r[o++] = mu_eff;
      r[o++] = G.hmax[c];
      r[o++] = G.pdyn[c];
      r[o++] = G.mvel[c];
      r[o++] = inertial;
    }
    if (what & IMEX_SNAP_Q) {
#pragma unroll
      for (int i = 0; i < NV; i++) sq[(size_t)c * NV + i] = q[i];
    }
    if (what & IMEX_SNAP_FIELDS) {
#pragma unroll
      for (int i = 0; i < NS; i++) sfld[(size_t)i * N + c] = dep[i];
      sfld[(size_t)NS * N + c] = G.hmax[c];
      sfld[(size_t)(NS + 1) * N + c] = G.pdyn[c];
      sfld[(size_t)(NS + 2) * N + c] = G.mvel[c];
#pragma unroll
      for (int i = 0; i < NS; i++) {
        sfld[(size_t)(NS + 3 + i) * N + c] = ero[i];
        sfld[(size_t)(2 * NS + 3 + i) * N + c] = erd[i];
      }
      sfld[(size_t)(3 * NS + 3) * N + c] = bcent ? bcent[c] : 0.0;
    }
  }
}

}  // namespace

// values per cell of each snapshot kind, in the order of the bits
static int snap_cols(const imex_handle* h, int kind) {
  const int nv = h->P.nv, ns = h->P.ns;
  switch (kind) {
    case 0: return nv;
    case 1: return 2 + nv;
    case 2: return 2 + ns;
    case 3: return p2d_cols(ns);
    case 4: return 3 * ns + 4;
    default: return 2 + ns;
  }
}
constexpr int NKIND = 7;

extern "C" int imex_snapshot_cols(imex_handle* h, unsigned kind) {
  if (!h) return -2;
  for (int b = 0; b < NKIND; b++)
    if (kind == (1u << b)) return snap_cols(h, b);
  return -2;
}

// the output kernels are instantiated for the n_solid + 4 variables of runs
// without pore pressure or the transported noise
static int no_pore(imex_handle* h, const char* what) {
  if (h->P.general) {
    h->err = std::string(what) + ": not built for runs with pore pressure or the transported noise";
    return -2;
  }
  return 0;
}

extern "C" int imex_snapshot_begin(imex_handle* h, unsigned what, double* q, double* q2d,
                                   double* dep2d, double* p2d, double* fields, double* ero2d,
                                   double* erd2d) {
  if (!h || (what & ~127u)) return -2;
  if (int rc = no_pore(h, "imex_snapshot_begin")) return rc;
  double* outs[NKIND] = {q, q2d, dep2d, p2d, fields, ero2d, erd2d};
  for (int b = 0; b < NKIND; b++)
    if ((what & (1u << b)) && !outs[b]) return -2;
  HCHK(h, hipSetDevice(h->device));
  if (!h->copy_stream) {
    HCHK(h, hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
    HCHK(h, hipEventCreateWithFlags(&h->ev_staged, hipEventDisableTiming));
    HCHK(h, hipEventCreateWithFlags(&h->ev_copied, hipEventDisableTiming));
  }
  const size_t N = (size_t)h->G.N;
  for (int b = 0; b < NKIND; b++) {
    if (!(what & (1u << b))) continue;
    const size_t need = N * snap_cols(h, b);
    if (h->stage_cap[b] < need) {
      if (h->copy_inflight) HCHK(h, hipEventSynchronize(h->ev_copied));
      HCHK(h, hipStreamSynchronize(h->stream));  // no earlier snapshot kernel still writes it
      if (int rc = dfree(h, h->stage[b])) return rc;
      h->stage[b] = nullptr;
      if (int rc = dalloc(h, &h->stage[b], need)) return rc;
      h->stage_cap[b] = need;
    }
  }
  // the staging of the previous snapshot may still be on its way to the host
  if (h->copy_inflight) HCHK(h, hipStreamWaitEvent(h->stream, h->ev_copied, 0));
  const int nb = std::max(1, std::min(8 * h->n_cu, (int)((N + 255) / 256)));
  const Params& P = h->P;
  const Grid& G = h->G;
#define SNAP_LAUNCH(NS)                                                                       \
  hipLaunchKernelGGL((h->fp32 ? k_snapshot<NS, float> : k_snapshot<NS, double>), dim3(nb),   \
                     dim3(256), 0, h->stream, P, G, h->G.bcent, what, h->stage[0], h->stage[1], \
                     h->stage[2], h->stage[3], h->stage[4], h->stage[5], h->stage[6])
  switch (P.ns) {
    case 1: SNAP_LAUNCH(1); break;
    case 2: SNAP_LAUNCH(2); break;
    default: SNAP_LAUNCH(3); break;
  }
#undef SNAP_LAUNCH
  HCHK(h, hipGetLastError());
  HCHK(h, hipEventRecord(h->ev_staged, h->stream));
  HCHK(h, hipStreamWaitEvent(h->copy_stream, h->ev_staged, 0));
  for (int b = 0; b < NKIND; b++)
    if (what & (1u << b))
      HCHK(h, hipMemcpyAsync(outs[b], h->stage[b], 8 * N * snap_cols(h, b), hipMemcpyDeviceToHost,
                             h->copy_stream));
  HCHK(h, hipEventRecord(h->ev_copied, h->copy_stream));
  h->copy_inflight = true;
  return 0;
}

extern "C" int imex_snapshot_wait(imex_handle* h) {
  if (!h) return -2;
  if (!h->copy_inflight) return 0;
  HCHK(h, hipSetDevice(h->device));
  HCHK(h, hipEventSynchronize(h->ev_copied));
  h->copy_inflight = false;
  return 0;
}

extern "C" int imex_host_alloc(size_t bytes, void** p) {
  if (!p) return -2;
  *p = nullptr;
  return hipHostMalloc(p, bytes ? bytes : 16, hipHostMallocDefault) == hipSuccess ? 0 : -1;
}

extern "C" int imex_host_free(void* p) {
  if (!p) return 0;
  return hipHostFree(p) == hipSuccess ? 0 : -1;
}

void imex_output_release(imex_handle* h) {
  if (h->copy_stream) {
    (void)hipStreamSynchronize(h->copy_stream);
    (void)hipStreamDestroy(h->copy_stream);
    h->copy_stream = nullptr;
  }
  if (h->ev_staged) (void)hipEventDestroy(h->ev_staged);
  if (h->ev_copied) (void)hipEventDestroy(h->ev_copied);
  h->ev_staged = h->ev_copied = nullptr;
}

extern "C" int imex_set_topography(imex_handle* h, const double* B_cent) {
  if (!h || !B_cent) return -2;
  HCHK(h, hipSetDevice(h->device));
  if (!h->G.bcent)
    if (int rc = dalloc(h, &h->G.bcent, (size_t)h->G.N)) return rc;
  HCHK(h, hipMemcpyAsync(h->G.bcent, B_cent, 8 * (size_t)h->G.N, hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  h->bcent_set = true;
  return 0;
}

extern "C" int imex_get_topography(imex_handle* h, double* B_cent) {
  if (!h || !B_cent) return -2;
  if (!h->G.bcent) {
    h->err = "imex_get_topography: B_cent was not set";
    return -2;
  }
  HCHK(h, hipSetDevice(h->device));
  HCHK(h, hipMemcpyAsync(B_cent, h->G.bcent, 8 * (size_t)h->G.N, hipMemcpyDeviceToHost, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int imex_runout_reduce(imex_handle* h, double x0_runout, double y0_runout,
                                  double alphas_threshold, unsigned flags, imex_runout_red* out) {
  if (!h || !out) return -2;
  if (int rc = no_pore(h, "imex_runout_reduce")) return rc;
  if ((flags & IMEX_RUNOUT_B) && !h->G.bcent) {
    h->err = "imex_runout_reduce: IMEX_RUNOUT_B needs imex_set_topography first";
    return -2;
  }
  HCHK(h, hipSetDevice(h->device));
  if (!h->hpos_old)
    if (int rc = dalloc(h, &h->hpos_old, (size_t)h->G.N)) return rc;
  const int nb = std::max(1, std::min(4 * h->n_cu, (h->G.N + RB - 1) / RB));
  if (h->red_cap < nb + 1) {
    RunoutPart* p = nullptr;
    if (int rc = dalloc(h, &p, (size_t)nb + 1)) return rc;
    h->d_red = p;
    h->red_cap = nb + 1;
  }
  RunoutPart* part = (RunoutPart*)h->d_red;
  const Params& P = h->P;
  const Grid& G = h->G;
#define RUNOUT_LAUNCH(NS)                                                                   \
  hipLaunchKernelGGL((h->fp32 ? k_runout<NS, float> : k_runout<NS, double>), dim3(nb),     \
                     dim3(RB), 0, h->stream, P, G, h->G.bcent, h->hpos_old, x0_runout,        \
                     y0_runout, alphas_threshold, flags, part)
  switch (P.ns) {
    case 1: RUNOUT_LAUNCH(1); break;
    case 2: RUNOUT_LAUNCH(2); break;
    default: RUNOUT_LAUNCH(3); break;
  }
#undef RUNOUT_LAUNCH
  HCHK(h, hipGetLastError());
  hipLaunchKernelGGL(k_runout_final, dim3(1), dim3(RB), 0, h->stream, part, nb);
  HCHK(h, hipGetLastError());
  RunoutPart r;
  HCHK(h, hipMemcpyAsync(&r, part + nb, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  out->max_h = r.maxh;
  out->sum_q1 = r.s[0] + r.e[0];
  out->sum_xq1 = r.s[1] + r.e[1];
  out->sum_yq1 = r.s[2] + r.e[2];
  out->n_hpos = (int64_t)r.nh;
  out->n_hpos_old = (int64_t)r.nold;
  out->n_new = (int64_t)r.nnew;
  out->imax = r.id;
  out->imax_x = r.idx;
  out->imax_y = r.idy;
  out->dist_max = r.d;
  out->dist_x_max = r.dxm;
  out->dist_y_max = r.dym;
  out->dist_at_imax_x = r.dax;
  out->dist_at_imax_y = r.day;
  out->ib = (flags & IMEX_RUNOUT_B) ? r.ib : -1;
  out->b_max = (flags & IMEX_RUNOUT_B) ? r.bm : 0.0;
  return 0;
}

extern "C" int imex_probes(imex_handle* h, int n, const double* xy, double* out) {
  if (!h || n < 0 || (n > 0 && (!xy || !out))) return -2;
  if (int rc = no_pore(h, "imex_probes")) return rc;
  if (n == 0) return 0;
  HCHK(h, hipSetDevice(h->device));
  const Params& P = h->P;
  const int nout = 6 + P.ns + (P.settling ? P.ns : 0);
  double *dxy = nullptr, *dout = nullptr;
  HCHK(h, hipMalloc(&dxy, 16 * (size_t)n));
  hipError_t e = hipMalloc(&dout, 8 * (size_t)n * nout);
  if (e != hipSuccess) {
    (void)hipFree(dxy);
    return set_err(h, "hipMalloc", e);
  }
  int rc = 0;
  e = hipMemcpyAsync(dxy, xy, 16 * (size_t)n, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    const dim3 grid((n + 63) / 64), blk(64);
#define PROBE_LAUNCH(NS)                                                                        \
  hipLaunchKernelGGL((h->fp32 ? k_probes<NS, float> : k_probes<NS, double>), grid, blk, 0, \
                     h->stream, P, h->G, n, dxy, dout)
    switch (P.ns) {
      case 1: PROBE_LAUNCH(1); break;
      case 2: PROBE_LAUNCH(2); break;
      default: PROBE_LAUNCH(3); break;
    }
#undef PROBE_LAUNCH
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, dout, 8 * (size_t)n * nout, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) rc = set_err(h, "imex_probes", e);
  (void)hipFree(dxy);
  (void)hipFree(dout);
  return rc;
}
