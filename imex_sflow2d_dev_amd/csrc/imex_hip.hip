// MI355X (gfx950) IMEX-RK stage of IMEX_SfloW2D: kernels and the C ABI.
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off -fPIC -shared
//        (see imex_sflow2d_dev_amd/build.py).  -ffp-contract=off keeps every
//        multiply and add separately rounded, as in the reference build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "imex_handle.h"

using namespace imex;

#define DINF __longlong_as_double(0x7ff0000000000000ULL)

// ===========================================================================
// active set: check_solve, src/solver_2d.f90:667-858
// ===========================================================================
// One 32x8 tile: seeds (frame, wet, source ring) of the tile and its two-cell
// halo in LDS, the diamond dilation of radius 2, the mask bytes, and the
// lists: each half of the tile (16 x 8 cells, entry 2 tile + half) goes to
// k_flux's list when it has active cells and is "hot" (a wet or ring cell
// within two cells of it), and a hot tile makes itself and its eight
// neighbours candidates of the next step's k_mask.
template <typename S>
__device__ __forceinline__ void mask_tile(const Params& P, const Grid& G, int tile, int solve_all,
                                          int par, unsigned int seq) {
  __shared__ unsigned char seed[RY][RX];
  __shared__ int s_any[2], s_cnt, s_hot[2];
  const int j0 = (tile % G.ntx) * TX, k0 = (tile / G.ntx) * TY;
  const int tid = threadIdx.x;
  __syncthreads();  // the previous tile's use of the shared flags is over
  if (tid == 0) {
    s_any[0] = s_any[1] = 0;
    s_cnt = 0;
    s_hot[0] = s_hot[1] = P.all_tiles | (int)G.st->hot;
  }
  __syncthreads();
  for (int e = tid; e < RX * RY; e += NTHR) {
    const int lj = e % RX, lk = e / RX;
    const int j = j0 + lj - HALO, k = k0 + lk - HALO;
    unsigned char s = 0;
    if (j >= 0 && j < G.nx && k >= 0 && k < G.ny) {
      const int c = j + G.nx * k;
      const bool ring = (j == 0 || j == G.nx - 1 || k == 0 || k == G.ny - 1);
      const bool wet = state_q<S>(G)[c] > S(0.0), src = P.radial && G.scell[c] == 2;
      if (ring || solve_all || wet || src) s = 1;
      // halo columns lj of half 0 (cells 0..FX-1): 0..FX+3; of half 1: FX..TX+3
      if ((wet || src) && lj < FX + 2 * HALO) s_hot[0] = 1;
      if ((wet || src) && lj >= FX) s_hot[1] = 1;
    }
    seed[lk][lj] = s;
  }
  __syncthreads();
  const int lj = tid % TX, lk = tid / TX;
  const int j = j0 + lj, k = k0 + lk;
  int act = 0;
  if (j < G.nx && k < G.ny) {
    const int c = j + G.nx * k;
    const bool ring = (j == 0 || j == G.nx - 1 || k == 0 || k == G.ny - 1);
    unsigned char m = 0;
    if (ring) {
      m = 1;
    } else {
#pragma unroll
      for (int dk = -2; dk <= 2; dk++) {
        const int r = 2 - (dk < 0 ? -dk : dk);
        for (int dj = -r; dj <= r; dj++) m |= seed[lk + HALO + dk][lj + HALO + dj];
      }
    }
    if (P.radial && G.scell[c] == 1) m = 0;
    G.mask[c] = m;
    act = m;
  }
  // active cells of the tile (one shared-memory atomic per wavefront) and of
  // each half
  const unsigned long long bal = __ballot(act != 0);
  if ((tid & 63) == 0 && bal) atomicAdd(&s_cnt, (int)__popcll(bal));
  if (act) s_any[lj >= FX ? 1 : 0] = 1;
  __syncthreads();
  const bool hot = s_hot[0] || s_hot[1];
  // the tile's global atomics, each on a thread of its own so that their
  // round trips overlap
  DevState* st = G.st;
  if (tid < 9 && hot) {
    // a hot tile makes itself and its neighbours candidates of the next pass;
    // the stamp lists each candidate once
    const int ux = tile % G.ntx + tid % 3 - 1, uy = tile / G.ntx + tid / 3 - 1;
    if (ux >= 0 && ux < G.ntx && uy >= 0 && uy < G.nty) {
      const int u = ux + G.ntx * uy;
      const unsigned int nslot = (seq + 1u) % 3u;
      if (atomicExch(&G.cand_stamp[u], seq + 1u) != seq + 1u) {
        const unsigned int slot = atomicAdd(&st->n_cand[nslot], 1u);
        G.cand[(size_t)nslot * G.ntx * G.nty + slot] = u;
      }
    }
  } else if ((tid == 9 || tid == 11) && s_any[tid == 11] && s_hot[tid == 11]) {
    // k_flux visits the half tiles with active cells that have a wet or
    // source-ring cell within two cells: elsewhere every face stencil is dry,
    // so every flux and wave speed is zero, and the (dry, all-zero) cells get
    // a zero right-hand side that k_update does not need
    const unsigned int slot = atomicAdd(&st->n_tiles[par], 1u);
    G.tile_list[(size_t)par * 2 * G.ntx * G.nty + slot] = 2 * tile + (tid == 11 ? 1 : 0);
  } else if (tid == 10) {
    // active cells of the tile beyond its frame band (counted by block 0)
    const long long extra = (long long)s_cnt - (long long)G.tile_static[tile];
    if (extra) atomicAdd(&st->active, (unsigned long long)extra);
  }
}

// full: every tile (the first pass after an upload, and after a pass with
// solve_all); otherwise the candidate tiles listed by the previous pass.  A
// tile outside the candidates had no wet or ring cell within ten cells at the
// previous step, so none within two now (fluxes move mass two cells at most
// per step): its mask bytes are the frame band's and stay as stored.
// The end of a device-driven step: dt history, t += dt, CFL minimum reset.
// The update launch of the step leaves it pending; the next launch on the
// stream, which no block of that update launch can overtake, applies it.
// A member of a batch stops once t reaches its t_end (the driver loop's
// condition t < t_end, src/IMEX_SfloW2D.f90:395).
__device__ __forceinline__ void commit_step(DevState* st) {
  if (st->pending) {
    st->dt_old_old = st->dt_old;
    st->dt_old = st->dt;
    st->t = st->t + st->dt;
    st->dtmin_bits = 0x7ff0000000000000ULL;
    st->pending = 0u;
    st->done = (st->t < st->t_end) ? 0u : 1u;
    st->steps = st->steps + 1u;
  }
}
__global__ void k_commit(Grid G) { commit_step(G.st); }

// driven: a step of the single-run device-side driver, which ends once t
// reaches t_end: the pass after the last step (and every later one of the
// call) commits that step and clears the slots, but builds no lists and
// counts no active cells; the driver's next call starts with a full pass.
template <typename S>
__device__ __forceinline__ void mask_body(const Params& P, const Grid& G, int solve_all, int par,
                                          unsigned int seq, int full, int driven = 0) {
  const bool stop = driven && G.st->stop_next;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    commit_step(G.st);
    // slots no block of this launch touches
    G.st->n_tiles[par ^ 1] = 0u;            // the next step's k_flux list
    G.st->n_work = 0u;                      // filled by the k_flux that follows
    G.st->n_cand[(seq + 2u) % 3u] = 0u;     // the list the next pass fills
    // the frame band of every tile; each visited tile adds what it has beyond
    if (!stop) atomicAdd(&G.st->active, G.static_active);
  }
  if (stop) return;
  if (full) {
    const int nt = G.ntx * G.nty;
    for (int b = blockIdx.x; b < nt; b += gridDim.x)
      mask_tile<S>(P, G, xcd_tile(b, nt), solve_all, par, seq);
  } else {
    const unsigned int slot = seq % 3u;
    const int n = (int)G.st->n_cand[slot];
    const int* list = G.cand + (size_t)slot * G.ntx * G.nty;
    for (int it = blockIdx.x; it < n; it += gridDim.x) mask_tile<S>(P, G, list[it], solve_all, par, seq);
  }
}
template <typename S>
__global__ __launch_bounds__(NTHR) void k_mask(Params P, Grid G, int solve_all, int par,
                                               unsigned int seq, int full, int driven) {
  mask_body<S>(P, G, solve_all, par, seq, full, driven);
}

// ===========================================================================
// reconstruction of one cell in one direction: src/solver_2d.f90:3095-3924
// ===========================================================================
// LDS primitive states of a half tile: sq[(i*RY + lk)*FRX + lj]
template <typename F>
struct TileView {
  const F* sq;
  const signed char* ssc;
  const unsigned char* ssd;
  bool tsvar;  // timestep variant: cells of the source disc carry T = T_ambient
  F T_amb;
  __device__ __forceinline__ F at(int i, int idx) const {
    F v = sq[i * FRX * RY + idx];
    if (tsvar && i == 3 && ssc[idx] == 1) v = T_amb;
    return v;
  }
};

template <typename F>
__device__ __forceinline__ void qp_to_qp2(const F* qp, F& u, F& v) {
  // src/constitutive_2d.f90:2164-2188 (qp2(1) = h + B is not needed here)
  if (qp[0] <= F(0.0)) {
    u = F(0.0);
    v = F(0.0);
  } else {
    u = qp[1] / qp[0];
    v = qp[2] / qp[0];
  }
}

// x direction, conservative-variable part: the W and E values of variable i
// of cell (j,k) at local index idx, and the stencil abscissae xs0 / xs2 the
// velocity part goes on with.
template <typename F>
__device__ __forceinline__ void recon_x_var(const Params& P, const Grid& G, const TileView<F>& T,
                                            int idx, int j, int i, F& Wi, F& Ei, F& xs0, F& xs2) {
  const int nx = G.nx;
  const bool src2 = (T.ssc[idx] == 2);
  const unsigned char sd = T.ssd[idx];
  const bool sE = sd & 1, sW = sd & 2;
  const F xs1 = F(G.xc[j]);
  const F v1 = T.at(i, idx);
  F pr;
  if (j == 0) {
    xs0 = F(G.xs[0]);
    xs2 = F(G.xc[1]);
    pr = src2 ? limit_minmod(v1, v1, T.at(i, idx + 1), xs0, xs1, xs2) : F(0.0);
  } else if (j == nx - 1) {
    xs2 = F(G.xs[nx]);
    xs0 = F(G.xc[nx - 2]);
    pr = src2 ? limit_minmod(T.at(i, idx - 1), v1, v1, xs0, xs1, xs2) : F(0.0);
  } else {
    xs0 = F(G.xc[j - 1]);
    xs2 = F(G.xc[j + 1]);
    F v0 = T.at(i, idx - 1), v2 = T.at(i, idx + 1);
    if (src2) {
      if (sE) {
        xs2 = F(G.xs[j + 1]);
        v2 = v1;
      } else if (sW) {
        xs0 = F(G.xs[j]);
        v0 = v1;
      }
    }
    pr = limit_minmod(v0, v1, v2, xs0, xs1, xs2);
  }
  const F dq = F(1.0) * F(P.dx2) * pr;
  Wi = v1 - dq;
  Ei = v1 + dq;
  if (j == 0) {
    if (src2)
      Wi = v1;
    else if (i == 1)
      Wi = fmin_f(Wi, F(0.0));
  }
  if (j == nx - 1) {
    if (src2)
      Ei = v1;
    else if (i == 1)
      Ei = fmax_f(Ei, F(0.0));
  }
}

// x direction: W and E face states (primitive) of cell (j,k) at local index idx.
template <int NS, typename F>
__device__ __forceinline__ void recon_x(const Params& P, const Grid& G, const TileView<F>& T, int idx,
                                        int j, int k, F* W, F* E) {
  constexpr int NV = 4 + NS;
  const int nx = G.nx;
  const signed char sc = T.ssc[idx];
  const unsigned char sd = T.ssd[idx];
  const bool src2 = (sc == 2);
  const bool sE = sd & 1, sW = sd & 2, sN = sd & 4, sS = sd & 8;
  const F xs1 = F(G.xc[j]);
  F xs0, xs2;
#pragma unroll
  for (int i = 0; i < NV; i++) recon_x_var<F>(P, G, T, idx, j, i, W[i], E[i], xs0, xs2);
  // u and v with the limiters of hu and hv; the y part of u may have stored
  // the north ring-face point in the x stencil (src/solver_2d.f90:3607)
  const bool typo = src2 && sN && !sS && (k > 0) && (k < G.ny - 1);
#pragma unroll
  for (int ii = 0; ii < 2; ii++) {
    const int i = NV + ii;
    const F v1 = T.at(i, idx);
    if (ii == 1 && typo) xs2 = F(G.ys[k + 1]);
    F v0, v2, q2u, q2v;
    if (j == 0) {
      qp_to_qp2(W, q2u, q2v);
      v0 = ii ? q2v : q2u;
      v2 = T.at(i, idx + 1);
    } else if (j == nx - 1) {
      qp_to_qp2(E, q2u, q2v);
      v0 = T.at(i, idx - 1);
      v2 = ii ? q2v : q2u;
    } else {
      v0 = T.at(i, idx - 1);
      v2 = T.at(i, idx + 1);
      if (src2) {
        if (sE) {
          xs2 = F(G.xs[j + 1]);
          v2 = v1;
        } else if (sW) {
          xs0 = F(G.xs[j]);
          v0 = v1;
        }
      }
    }
    const F pr = limit_minmod(v0, v1, v2, xs0, xs1, xs2);
    const F dq = F(1.0) * F(P.dx2) * pr;
    W[i] = v1 - dq;
    E[i] = v1 + dq;
    if (j == 0) {
      qp_to_qp2(W, q2u, q2v);
      W[i] = ii ? q2v : q2u;
    } else if (j == nx - 1) {
      qp_to_qp2(E, q2u, q2v);
      E[i] = ii ? q2v : q2u;
    } else if (src2) {
      if (sE) {
#pragma unroll
        for (int m = 0; m < NV + 2; m++) E[m] = T.at(m, idx);
      } else if (sW) {
#pragma unroll
        for (int m = 0; m < NV + 2; m++) W[m] = T.at(m, idx);
      }
    }
  }
}

// y direction: S and N face states of cell (j,k).
// y direction, conservative-variable part: the S and N values of variable i
// of cell (j,k), and the stencil ordinates the velocity part goes on with.
template <typename F>
__device__ __forceinline__ void recon_y_var(const Params& P, const Grid& G, const TileView<F>& T,
                                            int idx, int k, int i, F& Si, F& Ni, F& ys0, F& ys2) {
  const int ny = G.ny;
  const bool src2 = (T.ssc[idx] == 2);
  const unsigned char sd = T.ssd[idx];
  const bool sN = sd & 4, sS = sd & 8;
  const F ys1 = F(G.yc[k]);
  const F v1 = T.at(i, idx);
  F pr;
  if (k == 0) {
    ys0 = F(G.ys[0]);
    ys2 = F(G.yc[1]);
    pr = F(0.0);
  } else if (k == ny - 1) {
    ys0 = F(G.yc[ny - 2]);
    ys2 = F(G.ys[ny]);
    pr = F(0.0);
  } else {
    ys0 = F(G.yc[k - 1]);
    ys2 = F(G.yc[k + 1]);
    F v0 = T.at(i, idx - FRX), v2 = T.at(i, idx + FRX);
    if (src2) {
      if (sS) {
        ys0 = F(G.ys[k]);
        v0 = v1;
      } else if (sN) {
        ys2 = F(G.ys[k + 1]);
        v2 = v1;
      }
    }
    pr = limit_minmod(v0, v1, v2, ys0, ys1, ys2);
  }
  const F dq = F(1.0) * F(P.dy2) * pr;
  Si = v1 - dq;
  Ni = v1 + dq;
  if (k == 0 && i == 2) Si = fmin_f(Si, F(0.0));
  if (k == ny - 1 && i == 2) Ni = fmax_f(Ni, F(0.0));
}

template <int NS, typename F>
__device__ __forceinline__ void recon_y(const Params& P, const Grid& G, const TileView<F>& T, int idx,
                                        int j, int k, F* S, F* Nn) {
  constexpr int NV = 4 + NS;
  const int ny = G.ny;
  const signed char sc = T.ssc[idx];
  const unsigned char sd = T.ssd[idx];
  const bool src2 = (sc == 2);
  const bool sN = sd & 4, sS = sd & 8;
  const F ys1 = F(G.yc[k]);
  F ys0, ys2;
#pragma unroll
  for (int i = 0; i < NV; i++) recon_y_var<F>(P, G, T, idx, k, i, S[i], Nn[i], ys0, ys2);
#pragma unroll
  for (int ii = 0; ii < 2; ii++) {
    const int i = NV + ii;
    const F v1 = T.at(i, idx);
    F v0, v2, q2u, q2v;
    if (k == 0) {
      qp_to_qp2(S, q2u, q2v);
      v0 = ii ? q2v : q2u;
      v2 = T.at(i, idx + FRX);
    } else if (k == ny - 1) {
      qp_to_qp2(Nn, q2u, q2v);
      v0 = T.at(i, idx - FRX);
      v2 = ii ? q2v : q2u;
    } else {
      v0 = T.at(i, idx - FRX);
      v2 = T.at(i, idx + FRX);
      if (src2) {
        if (sS) {
          ys0 = F(G.ys[k]);
          v0 = v1;
        } else if (sN) {
          v2 = v1;
        }
      }
    }
    const F pr = limit_minmod(v0, v1, v2, ys0, ys1, ys2);
    const F dq = F(1.0) * F(P.dy2) * pr;
    S[i] = v1 - dq;
    Nn[i] = v1 + dq;
    if (k == 0) {
      qp_to_qp2(S, q2u, q2v);
      S[i] = ii ? q2v : q2u;
    } else if (k == ny - 1) {
      qp_to_qp2(Nn, q2u, q2v);
      Nn[i] = ii ? q2v : q2u;
    } else if (src2) {
      if (sS) {
#pragma unroll
        for (int m = 0; m < NV + 2; m++) S[m] = T.at(m, idx);
      } else if (sN) {
#pragma unroll
        for (int m = 0; m < NV + 2; m++) Nn[m] = T.at(m, idx);
      }
    }
  }
}

// The two states of one face from the primitive side states of its cells.
// dir 1: face between local cells idxA (west) and idxB (east); dir 2:
// south/north.  qpA / qpB hold on entry the reconstructed side states (E of
// A, W of B, or N of A, S of B) of the cells that exist (hasA / hasB); a side
// owned by a source-disc cell is the ring mirror when the ring cell on the
// other side faces the disc across this face, else the zero state (a face
// state the reference never writes).
template <int NS, typename F>
__device__ __forceinline__ void face_finish(const Params& P, const TileView<F>& T, int dir, int idxA,
                                            int idxB, bool hasA, bool hasB, F* qpA, F* qpB, F* qA,
                                            F* qB) {
  constexpr int NV = 4 + NS, NP = NV + 2;
  if (hasA) {
    if (T.ssc[idxA] == 1) {
#pragma unroll
      for (int i = 0; i < NP; i++) qpA[i] = F(0.0);
    }
    qp_to_qc<NS, F>(P, qpA, qA);
  }
  if (hasB) {
    if (T.ssc[idxB] == 1) {
#pragma unroll
      for (int i = 0; i < NP; i++) qpB[i] = F(0.0);
    }
    qp_to_qc<NS, F>(P, qpB, qB);
  }
  const int mom = dir;               // hu (1) or hv (2)
  const int vel = (dir == 1) ? NV : NV + 1;
  const unsigned char bitA = (dir == 1) ? 1 : 4;  // sourceE / sourceN of the A cell
  const unsigned char bitB = (dir == 1) ? 2 : 8;  // sourceW / sourceS of the B cell
  if (!hasA) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpA[i] = qpB[i];
#pragma unroll
    for (int i = 0; i < NV; i++) qA[i] = qB[i];
  } else if (!hasB) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpB[i] = qpA[i];
#pragma unroll
    for (int i = 0; i < NV; i++) qB[i] = qA[i];
  } else if (T.ssc[idxA] == 2 && (T.ssd[idxA] & bitA)) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpB[i] = qpA[i];
#pragma unroll
    for (int i = 0; i < NV; i++) qB[i] = qA[i];
    qB[mom] = -qA[mom];
    qpB[vel] = -qpA[vel];
  } else if (T.ssc[idxB] == 2 && (T.ssd[idxB] & bitB)) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpA[i] = qpB[i];
#pragma unroll
    for (int i = 0; i < NV; i++) qA[i] = qB[i];
    qA[mom] = -qB[mom];
    qpA[vel] = -qpB[vel];
  }
}

// eval_speeds for one face (the a- / a+ of kt_flux)
template <int NS, typename F>
__device__ __forceinline__ void kt_speeds(const Params& P, int dir, const F* qpA, const F* qpB,
                                          F& an, F& ap) {
  F amin, amax, bmin, bmax;
  local_speeds<NS, F>(P, qpA, dir, amin, amax);
  local_speeds<NS, F>(P, qpB, dir, bmin, bmax);
  an = fmin_f(fmin_f(amin, bmin), F(0.0));
  ap = fmax_f(fmax_f(amax, bmax), F(0.0));
}

// The wave speeds of one face in timestep's variant (cells of the source disc
// carry T = T_ambient): that variant's face states and kt_speeds, from the
// side states the tile already holds.  Under that view a side state differs only in its
// temperature (variable 3, next to a disc cell), which is recomputed here
// with the reconstruction's own rules; qpA / qpB hold on entry the side
// states in LDS (E or N of A, W or S of B).
template <int NS, typename F>
__device__ __forceinline__ void disc_speeds(const Params& P, const Grid& G, const TileView<F>& T0,
                                            int dir, int idxA, int idxB, int ja, int ka, int jb,
                                            int kb, bool hasA, bool hasB, F* qpA, F* qpB, F& an,
                                            F& ap) {
  constexpr int NV = 4 + NS, NP = NV + 2;
  F lo, hi, c0, c2;
  if (hasA) {
    if (T0.ssc[idxA] == 1) {
#pragma unroll
      for (int i = 0; i < NP; i++) qpA[i] = F(0.0);
    } else {
      const bool src2 = T0.ssc[idxA] == 2;
      const unsigned char sd = T0.ssd[idxA];
      if (dir == 1) {
        recon_x_var<F>(P, G, T0, idxA, ja, 3, lo, hi, c0, c2);
        // recon_x: a ring cell facing the disc eastwards keeps its own E state
        if (ja != 0 && ja != G.nx - 1 && src2 && (sd & 1)) hi = T0.at(3, idxA);
      } else {
        recon_y_var<F>(P, G, T0, idxA, ka, 3, lo, hi, c0, c2);
        // recon_y: ... northwards (and not southwards) keeps its own N state
        if (ka != 0 && ka != G.ny - 1 && src2 && !(sd & 8) && (sd & 4)) hi = T0.at(3, idxA);
      }
      qpA[3] = hi;
    }
  }
  if (hasB) {
    if (T0.ssc[idxB] == 1) {
#pragma unroll
      for (int i = 0; i < NP; i++) qpB[i] = F(0.0);
    } else {
      const bool src2 = T0.ssc[idxB] == 2;
      const unsigned char sd = T0.ssd[idxB];
      if (dir == 1) {
        recon_x_var<F>(P, G, T0, idxB, jb, 3, lo, hi, c0, c2);
        // ... westwards (and not eastwards) keeps its own W state
        if (jb != 0 && jb != G.nx - 1 && src2 && !(sd & 1) && (sd & 2)) lo = T0.at(3, idxB);
      } else {
        recon_y_var<F>(P, G, T0, idxB, kb, 3, lo, hi, c0, c2);
        // ... southwards keeps its own S state
        if (kb != 0 && kb != G.ny - 1 && src2 && (sd & 8)) lo = T0.at(3, idxB);
      }
      qpB[3] = lo;
    }
  }
  const int vel = (dir == 1) ? NV : NV + 1;
  const unsigned char bitA = (dir == 1) ? 1 : 4;
  const unsigned char bitB = (dir == 1) ? 2 : 8;
  if (!hasA) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpA[i] = qpB[i];
  } else if (!hasB) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpB[i] = qpA[i];
  } else if (T0.ssc[idxA] == 2 && (T0.ssd[idxA] & bitA)) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpB[i] = qpA[i];
    qpB[vel] = -qpA[vel];
  } else if (T0.ssc[idxB] == 2 && (T0.ssd[idxB] & bitB)) {
#pragma unroll
    for (int i = 0; i < NP; i++) qpA[i] = qpB[i];
    qpA[vel] = -qpB[vel];
  }
  kt_speeds<NS, F>(P, dir, qpA, qpB, an, ap);
}

// eval_speeds + eval_flux_KT + average_KT for one face:
// src/solver_2d.f90:3937-4014, :2834-3002, :3019-3049
template <int NS, typename F>
__device__ __forceinline__ void kt_flux(const Params& P, int dir, const F* qpA, const F* qpB,
                                        const F* qA, const F* qB, F* H, F& an, F& ap) {
  constexpr int NV = 4 + NS;
  F amin, amax, bmin, bmax;
  local_speeds<NS, F>(P, qpA, dir, amin, amax);
  local_speeds<NS, F>(P, qpB, dir, bmin, bmax);
  an = fmin_f(fmin_f(amin, bmin), F(0.0));
  ap = fmax_f(fmax_f(amax, bmax), F(0.0));
  F fA[NV], fB[NV];
  phys_flux<NS, F>(P, qA, qpA, dir, fA);
  phys_flux<NS, F>(P, qB, qpB, dir, fB);
#pragma unroll
  for (int i = 0; i < NV; i++) {
    if (an == ap) {
      H[i] = F(0.0);
    } else {
      const F avg = (ap * fA[i] - an * fB[i]) / (ap - an);
      H[i] = avg + (ap * an) / (ap - an) * (qB[i] - qA[i]);
    }
  }
  F S = F(0.0);
#pragma unroll
  for (int i = 0; i < NS; i++) S += H[4 + i];
  if (dir == 1) {
    if ((S / H[0]) >= F(1.0)) {
      F S2 = F(0.0);
#pragma unroll
      for (int i = 0; i < NS; i++) S2 += H[4 + i] / H[0];
#pragma unroll
      for (int i = 0; i < NS; i++) H[4 + i] = H[4 + i] / S2;
    }
    if ((qpA[1] == F(0.0)) && (qpB[1] == F(0.0))) {
      H[0] = F(0.0);
#pragma unroll
      for (int i = 3; i < NV; i++) H[i] = F(0.0);
    }
  } else {
    if ((S / H[0]) > F(1.0)) {
      const F S2 = S / H[0];
#pragma unroll
      for (int i = 0; i < NS; i++) H[4 + i] = H[4 + i] / S2;
    }
    if ((qA[2] == F(0.0)) && (qB[2] == F(0.0))) {
      H[0] = F(0.0);
#pragma unroll
      for (int i = 3; i < NV; i++) H[i] = F(0.0);
    }
  }
}

// ===========================================================================
// k_flux: timestep (src/solver_2d.f90:873-1006) fused with the stage-1
// explicit terms and eval_hyperbolic_terms (src/solver_2d.f90:1299-1338,
// :2616-2693).  R = divFlux - expl_terms per active cell.
// mode 0: finisher stores dt_cfl only; mode 1: also the driver's clip and
// growth limit (src/IMEX_SfloW2D.f90:417-439) into st->dt.
// ===========================================================================
// side states of the half tile's cells, computed once and read by both faces
// of each cell: x (W, E) for columns -1..FX of the TY rows, y (S, N) for rows
// -1..TY of the FX columns
constexpr int NRX = (FX + 2) * TY;
constexpr int NRY = FX * (TY + 2);

// faces of the half tile: x faces (FX+1) x TY, then y faces FX x (TY+1)
constexpr int NFX = (FX + 1) * TY;
constexpr int NFY = FX * (TY + 1);

// Blocks of FLUX_THR threads (single runs) take each phase of a half tile in
// one pass: the halo states with the active bytes, the side states, the faces
// with the cells' explicit terms, then the cells.  Blocks of NTHR threads
// (batch members, where many members keep the CUs busy) take two passes for
// the side states and the faces.
constexpr int FLUX_THR = 512;
template <int THR>
constexpr bool flux_one_pass() {
  return THR >= NFX + NFY + FX * TY;
}

template <int NS, typename F = double, int THR = NTHR>
constexpr size_t flux_lds_bytes() {
  constexpr int NV = 4 + NS, NP = NV + 2;
  return sizeof(F) * ((size_t)NP * RY * FRX + (size_t)NV * (FX + 1) * TY +
                           (size_t)NV * FX * (TY + 1) + (size_t)(FX + 1) * TY +
                           (size_t)FX * (TY + 1) + 2 * (size_t)NP * (NRX + NRY) +
                           (flux_one_pass<THR>() ? (size_t)NV * FX * TY : 0)) +
         2 * RY * FRX + FX * TY;
}

// F: the arithmetic and storage type of the state, the side states and the
// fluxes (double on the FP64 path; float in the FP32 form, where thin
// layers, h < 1e-3 m, are converted to primitive variables in double)
// One half tile (list entry ht: tile ht / 2, columns FX (ht % 2) onward) with
// SUB threads starting at thread tid0 of the block and the LDS set at base:
// the primitive states of its cells and two-cell halo, the side states, the
// faces, then the cells (divergence minus the explicit terms into R, the CFL
// candidate, the work list).  Every thread of the block calls this the same
// number of times, valid or not (the barriers are the block's); an invalid
// call does no work.
template <int NS, typename F, int SUB>
__device__ __forceinline__ void flux_half(const Params& P, const Grid& G, F* base, int tid, int ht,
                                          bool valid, double t, double& my_dt) {
  constexpr int NV = 4 + NS, NP = NV + 2;
  F* sq = base;
  F* sHx = sq + NP * RY * FRX;
  F* sHy = sHx + NV * (FX + 1) * TY;
  F* sAx = sHy + NV * FX * (TY + 1);
  F* sAy = sAx + (FX + 1) * TY;
  F* sRW = sAy + FX * (TY + 1);
  F* sRE = sRW + NP * NRX;
  F* sRS = sRE + NP * NRX;
  F* sRN = sRS + NP * NRY;
  const F* gq = state_q<F>(G);
  F* gR = state_R<F>(G);
  F* sEx = sRN + NP * NRY;  // [NV][FX*TY] explicit terms (one-pass blocks)
  constexpr bool ONE = flux_one_pass<SUB>();
  signed char* ssc = (signed char*)(sEx + (ONE ? NV * FX * TY : 0));
  unsigned char* ssd = (unsigned char*)(ssc + RY * FRX);
  unsigned char* sact = ssd + RY * FRX;
  const int nx = G.nx, ny = G.ny, N = G.N;
  const int tile = valid ? ht >> 1 : 0;
  const int j0 = (tile % G.ntx) * TX + (ht & 1) * FX, k0 = (tile / G.ntx) * TY;
  TileView<F> T1{sq, ssc, ssd, false, F(P.T_amb)};
  TileView<F> T0{sq, ssc, ssd, true, F(P.T_amb)};
  __syncthreads();  // the previous half tile's LDS reads are done
  if (valid) {
    // ---- primitive states of the tile and its two-cell halo ----------------
    for (int e = tid; e < FRX * RY; e += SUB) {
      const int lj = e % FRX, lk = e / FRX;
      const int j = j0 + lj - HALO, k = k0 + lk - HALO;
      F qp[NP];
      signed char sc = 0;
      unsigned char sd = 0;
#pragma unroll
      for (int i = 0; i < NP; i++) qp[i] = F(0.0);
      if (j >= 0 && j < nx && k >= 0 && k < ny) {
        const int c = j + nx * k;
        sc = G.scell[c];
        sd = G.sides[c];
        if (sc != 1) {
          F q[NV];
#pragma unroll
          for (int i = 0; i < NV; i++) q[i] = gq[(size_t)i * N + c];
          if (q[0] > F(0.0)) {
            F pd;
            qc_to_qp<NS, F>(P, q, qp, pd);
            if constexpr (sizeof(F) < sizeof(double)) {
              if (qp[0] < F(1.0e-3)) {  // thin layer: the division by h in double
                double qd[NV], qpd[NP], pdd;
#pragma unroll
                for (int i = 0; i < NV; i++) qd[i] = (double)q[i];
                qc_to_qp<NS, double>(P, qd, qpd, pdd);
#pragma unroll
                for (int i = 0; i < NP; i++) qp[i] = F(qpd[i]);
              }
            }
          } else {
            qp[3] = F(P.T_amb);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NP; i++) sq[i * FRX * RY + e] = qp[i];
      ssc[e] = sc;
      ssd[e] = sd;
    }
    // (one-pass blocks: the threads past the halo cells)
    for (int e = ONE ? tid - FRX * RY : tid; e >= 0 && e < FX * TY; e += SUB) {
      const int j = j0 + e % FX, k = k0 + e / FX;
      sact[e] = (j < nx && k < ny) ? G.mask[j + nx * k] : 0;
    }
  }
  __syncthreads();
  if (valid) {
    // ---- side states of the cells (cells of the source disc have none) -----
    for (int e = tid; e < NRX + NRY; e += SUB) {
      if (e < NRX) {
        const int lj = e % (FX + 2) - 1, lk = e / (FX + 2);
        const int j = j0 + lj, k = k0 + lk;
        const int idx = (lk + HALO) * FRX + lj + HALO;
        if (j >= 0 && j < nx && k < ny && ssc[idx] != 1) {
          F W[NP], E[NP];
          recon_x<NS, F>(P, G, T1, idx, j, k, W, E);
#pragma unroll
          for (int i = 0; i < NP; i++) {
            sRW[i * NRX + e] = W[i];
            sRE[i * NRX + e] = E[i];
          }
        }
      } else {
        const int e2 = e - NRX;
        const int lj = e2 % FX, lk = e2 / FX - 1;
        const int j = j0 + lj, k = k0 + lk;
        const int idx = (lk + HALO) * FRX + lj + HALO;
        if (k >= 0 && k < ny && j < nx && ssc[idx] != 1) {
          F S[NP], Nn[NP];
          recon_y<NS, F>(P, G, T1, idx, j, k, S, Nn);
#pragma unroll
          for (int i = 0; i < NP; i++) {
            sRS[i * NRY + e2] = S[i];
            sRN[i * NRY + e2] = Nn[i];
          }
        }
      }
    }
  }
  __syncthreads();
  if (valid) {
    // ---- faces: x faces (FX+1) x TY, then y faces FX x (TY+1) -------------
    // (one-pass blocks: the threads past the faces evaluate the explicit
    // terms of the cells, which depend on the cell's own state only)
    for (int e0 = tid; e0 < NFX + NFY + (ONE ? FX * TY : 0); e0 += SUB) {
     if (ONE && e0 >= NFX + NFY) {
      const int ce = e0 - NFX - NFY;
      const int lj = ce % FX, lk = ce / FX;
      const int j = j0 + lj, k = k0 + lk;
      if (j < nx && k < ny && sact[ce]) {
        const int c = j + nx * k;
        const int idx = (lk + HALO) * FRX + lj + HALO;
        F qp[NP], ex[NV];
#pragma unroll
        for (int i = 0; i < NP; i++) qp[i] = sq[i * FRX * RY + idx];
        F ap = F(0.0), ax = F(0.0), ay = F(0.0);
        if (ssc[idx] == 2) {
          ap = F(G.arcp[c]);
          ax = F(G.arcx[c]);
          ay = F(G.arcy[c]);
        }
        expl_terms<NS, F>(P, F(G.bx[c]), F(G.by[c]), qp, ex, t, ap, ax, ay);
#pragma unroll
        for (int i = 0; i < NV; i++) sEx[i * FX * TY + ce] = ex[i];
      }
     } else if (e0 < NFX) {
      const int e = e0;
      const int fl = e % (FX + 1), fk = e / (FX + 1);
      const int jf = j0 + fl, k = k0 + fk;
      bool need = (k < ny) && (jf <= nx) &&
                  ((fl > 0 && sact[fk * FX + fl - 1]) || (fl < FX && jf < nx && sact[fk * FX + fl]));
      // a face whose four stencil cells are dry has zero states on both
      // sides: H = 0 and no wave speed (the minmod slope of a dry cell is 0)
      if (need) {
        const int b = (fk + HALO) * FRX + fl + HALO;
        need = (sq[b - 2] != F(0.0)) || (sq[b - 1] != F(0.0)) || (sq[b] != F(0.0)) ||
               (sq[b + 1] != F(0.0));
      }
      F H[NV];
      F amax = F(0.0);
#pragma unroll
      for (int i = 0; i < NV; i++) H[i] = F(0.0);
      if (need) {
        const int lk = fk + HALO, ljb = fl + HALO;
        const int idxB = lk * FRX + ljb, idxA = idxB - 1;
        const bool hasA = jf - 1 >= 0, hasB = jf < nx;
        const int ra = fk * (FX + 2) + fl;  // side states of cell A; B is ra + 1
        F qpA[NP], qpB[NP], qA[NV], qB[NV], an, ap;
#pragma unroll
        for (int i = 0; i < NP; i++) {
          qpA[i] = sRE[i * NRX + ra];
          qpB[i] = sRW[i * NRX + ra + 1];
        }
        face_finish<NS, F>(P, T1, 1, idxA, idxB, hasA, hasB, qpA, qpB, qA, qB);
        kt_flux<NS, F>(P, 1, qpA, qpB, qA, qB, H, an, ap);
        bool disc = false;
        if (P.radial) {
          for (int d = -2; d <= 1; d++) {
            const int jj = jf + d;
            if (jj >= 0 && jj < nx && ssc[lk * FRX + ljb + d] == 1) disc = true;
          }
        }
        if (disc) {  // timestep's wave speeds see T_ambient in the source disc
#pragma unroll
          for (int i = 0; i < NP; i++) {
            qpA[i] = sRE[i * NRX + ra];
            qpB[i] = sRW[i * NRX + ra + 1];
          }
          disc_speeds<NS, F>(P, G, T0, 1, idxA, idxB, jf - 1, k, jf, k, hasA, hasB, qpA, qpB, an, ap);
        }
        amax = fmax_f(ap, -an);
      }
#pragma unroll
      for (int i = 0; i < NV; i++) sHx[i * (FX + 1) * TY + e] = H[i];
      sAx[e] = amax;
     } else {
      const int e = e0 - NFX;
      const int fj = e % FX, fl = e / FX;
      const int j = j0 + fj, kf = k0 + fl;
      bool need = (j < nx) && (kf <= ny) &&
                  ((fl > 0 && sact[(fl - 1) * FX + fj]) || (fl < TY && kf < ny && sact[fl * FX + fj]));
      if (need) {
        const int b = (fl + HALO) * FRX + fj + HALO;
        need = (sq[b - 2 * FRX] != F(0.0)) || (sq[b - FRX] != F(0.0)) || (sq[b] != F(0.0)) ||
               (sq[b + FRX] != F(0.0));
      }
      F H[NV];
      F amax = F(0.0);
#pragma unroll
      for (int i = 0; i < NV; i++) H[i] = F(0.0);
      if (need) {
        const int lj = fj + HALO, lkb = fl + HALO;
        const int idxB = lkb * FRX + lj, idxA = idxB - FRX;
        const bool hasA = kf - 1 >= 0, hasB = kf < ny;
        const int ra = fl * FX + fj;  // side states of cell A; B is ra + FX
        F qpA[NP], qpB[NP], qA[NV], qB[NV], an, ap;
#pragma unroll
        for (int i = 0; i < NP; i++) {
          qpA[i] = sRN[i * NRY + ra];
          qpB[i] = sRS[i * NRY + ra + FX];
        }
        face_finish<NS, F>(P, T1, 2, idxA, idxB, hasA, hasB, qpA, qpB, qA, qB);
        kt_flux<NS, F>(P, 2, qpA, qpB, qA, qB, H, an, ap);
        bool disc = false;
        if (P.radial) {
          for (int d = -2; d <= 1; d++) {
            const int kk = kf + d;
            if (kk >= 0 && kk < ny && ssc[(lkb + d) * FRX + lj] == 1) disc = true;
          }
        }
        if (disc) {  // timestep's wave speeds see T_ambient in the source disc
#pragma unroll
          for (int i = 0; i < NP; i++) {
            qpA[i] = sRN[i * NRY + ra];
            qpB[i] = sRS[i * NRY + ra + FX];
          }
          disc_speeds<NS, F>(P, G, T0, 2, idxA, idxB, j, kf - 1, j, kf, hasA, hasB, qpA, qpB, an, ap);
        }
        amax = fmax_f(ap, -an);
      }
#pragma unroll
      for (int i = 0; i < NV; i++) sHy[i * FX * (TY + 1) + e] = H[i];
      sAy[e] = amax;
     }
    }
  }
  __syncthreads();
  if (valid) {
    // ---- cells: divergence, explicit terms, CFL candidate -----------------
    {
      const int lj = tid % FX, lk = tid / FX;
      const int j = j0 + lj, k = k0 + lk;
      bool work = false;
      const int c = j + nx * k;
      if (tid < FX * TY && j < nx && k < ny && sact[tid]) {
        const int ew = lk * (FX + 1) + lj;
        const int es = lk * FX + lj;
        F ex[NV];
        if constexpr (ONE) {
#pragma unroll
          for (int i = 0; i < NV; i++) ex[i] = sEx[i * FX * TY + tid];
        } else {
          F qp[NP];
          const int idx = (lk + HALO) * FRX + lj + HALO;
#pragma unroll
          for (int i = 0; i < NP; i++) qp[i] = sq[i * FRX * RY + idx];
          const signed char sc = ssc[idx];
          F ap = F(0.0), ax = F(0.0), ay = F(0.0);
          if (sc == 2) {
            ap = F(G.arcp[c]);
            ax = F(G.arcx[c]);
            ay = F(G.arcy[c]);
          }
          expl_terms<NS, F>(P, F(G.bx[c]), F(G.by[c]), qp, ex, t, ap, ax, ay);
        }
#pragma unroll
        for (int i = 0; i < NV; i++) {
          F d = F(0.0);
          d = d + (sHx[i * (FX + 1) * TY + ew + 1] - sHx[i * (FX + 1) * TY + ew]) * F(P.one_by_dx);
          d = d + (sHy[i * FX * (TY + 1) + es + FX] - sHy[i * FX * (TY + 1) + es]) * F(P.one_by_dy);
          const F r = d - ex[i];
          gR[(size_t)i * N + c] = r;
          work = work || (r != F(0.0));
        }
        // A cell whose state and right-hand side are all zero stays exactly
        // zero through stage 2, the assembly, sedimentation and the epilogue,
        // and raises no error unless T_ambient < 273 K: only the other cells
        // go to k_update.
#pragma unroll
        for (int i = 0; i < NV; i++) work = work || (gq[(size_t)i * N + c] != F(0.0));
        work = work || (P.T_amb < 273.0);
        double ma = (double)fmax_f(sAx[ew], sAx[ew + 1]);
        const double dtx = (ma > 0.0) ? P.cfl * P.dx / ma : DINF;
        ma = (double)fmax_f(sAy[es], sAy[es + FX]);
        const double dty = (ma > 0.0) ? P.cfl * P.dy / ma : DINF;
        const double dtc = fmin_f(dtx, dty);
        my_dt = (dtc < my_dt) ? dtc : my_dt;
      }
      // compact the work cells: one atomic per wavefront
      const unsigned long long bal = __ballot(work);
      const int lane = tid & 63;
      unsigned int base = 0u;
      if (lane == 0 && bal) base = atomicAdd(&G.st->n_work, (unsigned int)__popcll(bal));
      base = __shfl(base, 0, 64);
      if (work) G.work[base + __popcll(bal & ((1ULL << lane) - 1ULL))] = c;
    }
  
  }
}

// k_flux's work: the hot half tiles of the latest active-set pass.  THR =
// NTHR: each block takes half tiles one after the other.  THR = 2 NTHR (one
// block per CU, the grid no larger than the CUs): with at most one half tile
// per block, all THR threads take it and each phase is one pass; with more
// half tiles, the two NTHR-thread halves of a block each take one, in their
// own LDS set.  The results do not depend on the split.
template <int NS, typename F = double, int THR = NTHR>
__device__ __forceinline__ void flux_body(const Params& P, const Grid& G, int mode, double t_arg,
                                          int par) {
  extern __shared__ double smem[];
  __shared__ double red[THR / 64];
  const int tid = threadIdx.x;
  double my_dt = DINF;
  const int n_list = (int)G.st->n_tiles[par];
  const int* list = G.tile_list + (size_t)par * 2 * G.ntx * G.nty;
  const double t = (mode == 1) ? G.st->t : t_arg;
  F* base = reinterpret_cast<F*>(smem);
  if constexpr (THR == NTHR) {
    for (int it = blockIdx.x; it < n_list; it += gridDim.x)
      flux_half<NS, F, NTHR>(P, G, base, tid, list[it], true, t, my_dt);
  } else {
    static_assert(THR == 2 * NTHR, "k_flux blocks are NTHR or 2 NTHR threads");
    if (n_list <= (int)gridDim.x) {
      for (int it = blockIdx.x; it < n_list; it += gridDim.x)
        flux_half<NS, F, THR>(P, G, base, tid, list[it], true, t, my_dt);
    } else {
      const int sub = tid / NTHR;
      F* mine = base + sub * (flux_lds_bytes<NS, F, NTHR>() / sizeof(F));
      for (int pp = blockIdx.x; 2 * pp < n_list; pp += gridDim.x) {
        const int it = 2 * pp + sub;
        flux_half<NS, F, NTHR>(P, G, mine, tid - sub * NTHR, (it < n_list) ? list[it] : 0,
                               it < n_list, t, my_dt);
      }
    }
  }
  // ---- CFL min: wavefront shuffle, LDS, one atomic per block ---------------
  // (every candidate is positive, so the IEEE bit patterns order as the values)
  double v = my_dt;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    double b = red[0];
    for (int w = 1; w < THR / 64; w++) b = (red[w] < b) ? red[w] : b;
    if (b < DINF) atomicMin(&G.st->dtmin_bits, (unsigned long long)__double_as_longlong(b));
  }
}

template <int NS, typename F>
__global__ __launch_bounds__(FLUX_THR) void k_flux(Params P, Grid G, int mode, double t_arg,
                                                   int par) {
  flux_body<NS, F, FLUX_THR>(P, G, mode, t_arg, par);
}

// The CFL time step from the blocks' minimum (timestep's result,
// src/solver_2d.f90:873-1006), for the host-driven step.  One thread; runs
// after k_flux, so the kernel boundary orders it after every block's atomic.
__global__ void k_dt(Params P, Grid G) {
  DevState* st = G.st;
  st->dt_cfl = fmin_f(P.max_dt, __longlong_as_double((long long)st->dtmin_bits));
  st->dtmin_bits = 0x7ff0000000000000ULL;
}

// (driver_dt, the device-side driver's dt, is in imex_handle.h; it reads
// only fields no kernel of the step writes before the sedimentation launch)

// ===========================================================================
// stage 2 of imex_RK_solver for one cell: src/solver_2d.f90:1103-1338 with
// solve_rk_step / lnsrch / eval_f / eval_jacobian (:1662-2367) and the
// assembly checks (:1342-1630)
// ===========================================================================
template <int NS>
__device__ void eval_f(const Params& P, double xi, double dt, const double* qj, const double* q0,
                       const double* Rni, double* f, double& scal, PhaseClk* pc = nullptr) {
  constexpr int NV = 4 + NS;
  double n1, n2;
  if (!implicit_r<NS>(P, qj, n1, n2, xi)) {
    pc_add(pc, PC_CPLX, 1ULL);
    cx qc[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) qc[i] = C(qj[i], 0.0);
    cx c1, c2;
    implicit_c<NS>(P, qc, c1, c2, xi);
    n1 = c1.re;
    n2 = c2.re;
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double nh = (i == 1) ? n1 : ((i == 2) ? n2 : 0.0);
    const double Rj = Rni[i] - 1.0 * nh;
    f[i] = qj[i] - q0[i] + dt * Rj;
    f[i] = 1.0 * f[i];
  }
#pragma unroll
  for (int i = 0; i < NV; i++) s += f[i] * f[i];
  scal = 0.5 * s;
}

// Jacobian: diagonal qorg for the explicit rows, the 2x2 momentum block from
// the complex step (all other entries are zero).  LANES = 2: lane `lane` of an
// aligned pair evaluates column 1 + lane and the pair swaps the entries, so
// the two complex evaluations run side by side; both lanes call together.
template <int NS, int LANES>
__device__ void eval_jacobian(const Params& P, double xi, double dt, const double* qrel,
                              const double* qorg,
                              double& L11, double& L21, double& L12, double& L22, int lane) {
  constexpr int NV = 4 + NS;
#pragma unroll
  for (int k = 0; k < 3 - LANES; k++) {
    const int col = (LANES == 2) ? 1 + lane : 1 + k;
    cx qc[NV];
#pragma unroll
    for (int m = 0; m < NV; m++)
      qc[m] = cmul(C(qrel[m], (m == col) ? P.hcs : 0.0), C(qorg[m], 0.0));
    cx n1, n2;
    implicit_c<NS>(P, qc, n1, n2, xi);
    const double J1 = 1.0 * n1.im * P.one_by_hcs;
    const double J2 = 1.0 * n2.im * P.one_by_hcs;
    // column col: (L1col, L2col) = (diag_1 - dt J1, diag_2 - dt J2) with the
    // diagonal qorg entry in row col and 0 in the other row
    const double a = ((col == 1) ? (1.0 * qorg[1]) : 0.0) - dt * 1.0 * J1;
    const double b = ((col == 1) ? 0.0 : (1.0 * qorg[2])) - dt * 1.0 * J2;
    if constexpr (LANES == 2) {
      // selects rather than stores into a lane-indexed column: the latter
      // puts the four entries in scratch memory
      const double oa = __shfl_xor(a, 1), ob = __shfl_xor(b, 1);
      const bool first = lane == 0;
      L11 = first ? a : oa;
      L21 = first ? b : ob;
      L12 = first ? oa : a;
      L22 = first ? ob : b;
    } else if (col == 1) {
      L11 = a;
      L21 = b;
    } else {
      L12 = a;
      L22 = b;
    }
  }
}

// column entry L(r, c) of the sparse Jacobian
__device__ __forceinline__ double Lrc(int r, int c, const double* qorg, double L11, double L21,
                                      double L12, double L22) {
  if (c == 1) return (r == 1) ? L11 : ((r == 2) ? L21 : 0.0);
  if (c == 2) return (r == 1) ? L12 : ((r == 2) ? L22 : 0.0);
  return (r == c) ? 1.0 * qorg[c] : 0.0;
}

// lnsrch, src/solver_2d.f90:2021-2226.  Returns true when rt and f_out are
// the residual at the returned x (the caller then need not evaluate it
// again: eval_f is a pure function).
//
// The reference's step sequence: alam_0 = 1; after a trial at alam that
// neither accepts (f <= 0.9 f_old) nor ends the search (alam < alamin), the
// next alam is MAX(tmplam, alam/2), where tmplam is the quadratic minimiser
// when alam == 1 and otherwise a cubic minimiser that the reference first
// caps at alam/2 -- so every trial after one at alam != 1 is at exactly half
// the previous alam (a NaN tmplam loses both comparisons and gives alam/2
// too).  The trials of a search are thus known in advance except after a
// trial at alam == 1.  LANES = 2: the lanes of the pair evaluate the next two
// trials of the sequence at once, the second assuming the halving; the pair
// then walks the two results in order with the reference's rules and keeps
// the second only when its alam is the one those rules give.  Extra
// evaluations past the deciding trial are discarded.
template <int NS, int LANES>
__device__ bool lnsrch(const Params& P, double xi, double dt, const double* xold, const double* qorg,
                       const double* q0, double fold, const double* g, double* p, double* x,
                       double& f_out, double* rt, double stpmax, const double* Rni,
                       unsigned int& err, int lane, PhaseClk* pc = nullptr) {
  constexpr int NV = 4 + NS;
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
    return true;  // rt and f_out still hold the residual at xold
  }
  // value of the pair's lane l (the own value when LANES == 1)
  auto from = [&](double v, int l) {
    if constexpr (LANES == 2) {
      const double o = __shfl_xor(v, 1);
      return (l == lane) ? v : o;
    } else {
      return v;
    }
  };
  double base = 1.0;  // alam of the round's first trial
  double xt[NV], qj[NV], rtt[NV], ft = 0.0;
  // The reference loop ends by construction: alam at least halves after the
  // first trial, so it drops below any positive alamin within
  // log2(DBL_MAX / DBL_TRUE_MIN) < 2100 trials.  Only a NaN alamin (a NaN
  // step) would loop for ever; the bound turns that into an error.
  int trials = 0;
  while (true) {
    pc_add(pc, PC_TRIALS, 1ULL);
    const double a = (lane == 0) ? base : 0.5 * base;
#pragma unroll
    for (int i = 0; i < NV; i++) xt[i] = xold[i] + a * p[i];
#pragma unroll
    for (int i = 0; i < NV; i++) qj[i] = xt[i] * qorg[i];
    eval_f<NS>(P, xi, dt, qj, q0, Rni, rtt, ft, pc);
    double next = base;
    int decided = -1;
    bool accept = false;
#pragma unroll
    for (int l = 0; l < LANES; l++) {
      const double al = from(a, l), fl = from(ft, l);
      if (al != next) break;  // the halving guess of lane l does not hold
      trials++;
      if (fl <= 0.9 * fold) {
        decided = l;
        accept = true;
        break;
      }
      if (al < alamin) {
        decided = l;
        break;
      }
      next = (al == 1.0) ? fmax_f(-(slope / (2.0 * (fl - fold - slope))), 0.5 * al) : 0.5 * al;
    }
    if (decided >= 0) {
      const double al = from(a, decided);
#pragma unroll
      for (int i = 0; i < NV; i++) rt[i] = from(rtt[i], decided);
      if (accept) {
#pragma unroll
        for (int i = 0; i < NV; i++) x[i] = xold[i] + al * p[i];
        f_out = from(ft, decided);
        return true;
      }
#pragma unroll
      for (int i = 0; i < NV; i++) x[i] = xold[i];
      f_out = fold;
      return false;  // rt is the last trial's residual
    }
    if (trials >= 2200) {  // lane 0's trial of the round, so that the pair stays in step
      err |= IMEX_ERR_NAN;
#pragma unroll
      for (int i = 0; i < NV; i++) {
        x[i] = from(xt[i], 0);
        rt[i] = from(rtt[i], 0);
      }
      f_out = from(ft, 0);
      return false;
    }
    base = next;
  }
}

template <int NS, int LANES>
__device__ int solve_rk_step(const Params& P, double xi, double dt, double* qj, const double* q0,
                             const double* Rni, unsigned int& err, int lane,
                             PhaseClk* pc = nullptr) {
  constexpr int NV = 4 + NS;
  const double TOLF = 1.0e-10, STPMX = 100.0;
  double qorg[NV], qrel[NV], rt[NV], xold[NV], p[NV], grad[NV];
  double scal_f, scal_f_old;
  double L11 = 0.0, L21 = 0.0, L12 = 0.0, L22 = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) grad[i] = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) qorg[i] = fmax_f(fabs(qj[i]), 1.0e-3);
#pragma unroll
  for (int i = 0; i < NV; i++) qrel[i] = qj[i] / qorg[i];
  int it;
  bool fresh = false;  // rt and scal_f already hold the residual at qj
  for (it = 1; it <= 100; it++) {
    if (!fresh) eval_f<NS>(P, xi, dt, qj, q0, Rni, rt, scal_f, pc);
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mx = fmax_f(mx, fabs(rt[i]));
    if (mx < TOLF) break;
    const unsigned long long cj = pc_now();
    eval_jacobian<NS, LANES>(P, xi, dt, qrel, qorg, L11, L21, L12, L22, lane);
    pc_add(pc, PC_JAC, pc_now() - cj);
    double d1[NV - 2];
    int ne = 0;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      if (r == 1 || r == 2) continue;
      d1[ne] = rt[r] / (1.0 * qorg[r]);
      ne++;
    }
    double d2a = rt[1], d2b = rt[2];
    {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int a = 0; a < NV - 2; a++) s0 += d1[a] * 0.0;
#pragma unroll
      for (int a = 0; a < NV - 2; a++) s1 += d1[a] * 0.0;
      d2a = d2a - s0;
      d2b = d2b - s1;
    }
    const double inv_det = 1.0 / (L11 * L22 - L21 * L12);
    const double s1 = (d2a * L22 - d2b * L12) * inv_det;
    const double s2 = (L11 * d2b - L21 * d2a) * inv_det;
    ne = 0;
#pragma unroll
    for (int r = 0; r < NV; r++) {
      if (r == 1)
        p[r] = -s1 + 0.0;
      else if (r == 2)
        p[r] = -s2 + 0.0;
      else {
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
        for (int r = 0; r < NV; r++) s += rt[r] * Lrc(r, i, qorg, L11, L21, L12, L22);
        grad[i] = s;
      }
      const unsigned long long cl = pc_now();
      fresh = lnsrch<NS, LANES>(P, xi, dt, xold, qorg, q0, scal_f_old, grad, p, qrel, scal_f, rt,
                                stpmax, Rni, err, lane, pc);
      pc_add(pc, PC_LNSRCH, pc_now() - cl);
    } else {
#pragma unroll
      for (int i = 0; i < NV; i++) qrel[i] = xold[i] + p[i];
#pragma unroll
      for (int i = 0; i < NV; i++) qj[i] = qrel[i] * qorg[i];
      eval_f<NS>(P, xi, dt, qj, q0, Rni, rt, scal_f, pc);
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
  return it;
}

// stage 2 + assembly for one cell; q holds q0 on entry, the new state on exit
template <int NS, int LANES>
__device__ void stage2(const Params& P, double xi, double dt, double* q, const double* X,
                       unsigned int& err, int lane, PhaseClk* pc = nullptr) {
  pc_mark(pc, PC_STAGE2);
  constexpr int NV = 4 + NS, NP = NV + 2;
  double q0[NV], qfv[NV], qp[NP], q_si[NV], qg[NV], si[NV], Rj[NV], pd;
#pragma unroll
  for (int i = 0; i < NV; i++) q0[i] = q[i];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double A = 0.0;
    A = A + X[i] * 1.0;
    A = A + (0.0 - 0.0) * 0.0;
    double B = 0.0;
    B = B + (0.0 + 0.0) * 0.0;
    B = B + (0.0 + 0.0) * 0.0;
    qfv[i] = q0[i] - dt * (A - B);
  }
  qc_to_qp<NS>(P, qfv, qp, pd);
  if (qfv[0] > 0.0) {
    semi_impl<NS>(P, qp, si);
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
      const double B = 0.0 + (0.0 + 0.0) * 0.0;
      Rj[i] = (A - B) - 1.0 * si[i];
    }
    pc_mark(pc, PC_NEWTON0);
    const int its = solve_rk_step<NS, LANES>(P, xi, dt, qg, q0, Rj, err, lane, pc);
    pc_mark(pc, PC_NEWTON1);
    pc_add(pc, PC_ITS, (unsigned long long)its);
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
  // assemble_sol
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
  qc_to_qp<NS>(P, q, qp, pd);
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

// update_erosion_deposition_cell for one cell: src/solver_2d.f90:2429-2593
// erod / ero: the cell's erodible layer and eroded thickness (updated; the
// caller passes zeros and ignores them when erosion is off)
template <int NS, int LANES>
__device__ void erosion_deposition(const Params& P, double dt, double* q, double* depo,
                                   double* erod, double* ero, double& topo, unsigned int& err,
                                   const ddm::dd* tab, int lane, PhaseClk* pc = nullptr) {
  constexpr int NV = 4 + NS, NP = NV + 2;
  double qp[NP], pd, dep[NS], eq[NV], ers[NS];
  if (q[0] > 0.0) {
    qc_to_qp<NS>(P, q, qp, pd);
  } else {
#pragma unroll
    for (int i = 0; i < NP; i++) qp[i] = 0.0;
    qp[3] = P.T_amb;
  }
  if (mass_exchange<NS, LANES>(P, qp, erod, dt, dep, eq, ers, topo, tab, lane, pc))
    err |= IMEX_ERR_NEG_DEP;
#pragma unroll
  for (int i = 0; i < NV; i++) q[i] = q[i] + dt * eq[i];
#pragma unroll
  for (int i = 0; i < NS; i++) q[4 + i] = fmax_f(0.0, q[4 + i]);
#pragma unroll
  for (int i = 0; i < NS; i++) depo[i] = depo[i] + dt * dep[i];
#pragma unroll
  for (int i = 0; i < NS; i++) ero[i] = ero[i] + dt * ers[i];
#pragma unroll
  for (int i = 0; i < NS; i++) erod[i] = erod[i] - dt * ers[i];
  if (P.erodible_deposit) {
#pragma unroll
    for (int i = 0; i < NS; i++) erod[i] = erod[i] + dt * dep[i];
  }
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
    qc_to_qp<NS>(P, q, qp, pd);
    rg = mixt_var<NS>(P, qp, false).red_grav;
  } else {
    rg = 0.0;
  }
  if (rg <= 0.0) {
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = 0.0;
  }
}

// driver epilogue for one cell: src/IMEX_SfloW2D.f90:449-505 (first: :306-353)
template <int NS>
__device__ void epilogue(const Params& P, const double* q, double& hmax, double& pdynmax,
                         double& mvel, unsigned long long& vuln, bool first) {
  constexpr int NV = 4 + NS, NP = NV + 2;
  if (q[0] > 0.0) {
    double qp[NP], pd;
    qc_to_qp<NS>(P, q, qp, pd);
    hmax = first ? qp[0] : fmax_f(hmax, qp[0]);
    const double ru = qp[NV], rv = qp[NV + 1];
    const double mv = sqrt(ru * ru + rv * rv);
    if (qp[0] > 0.001) {
      pdynmax = first ? pd : fmax_f(pdynmax, pd);
      mvel = first ? mv : fmax_f(mvel, mv);
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
    vuln = first ? bits : (vuln | bits);
  } else if (first) {
    hmax = 0.0;
    pdynmax = 0.0;
  }
}

// PARTS: 1 stage 2 + assembly, 2 erosion/deposition, 4 epilogue, 8 initial
// epilogue (over every cell).  Without 8 the kernel walks the work list the
// step's k_flux built.  mode 1: dt from st->dt, and the launch holding the
// epilogue advances t (device-side driver).  Each PARTS is its own kernel so
// that the register budget of the Newton solve does not bound the occupancy
// of the sedimentation and epilogue.
// mode 1 (device-side driver): the stage-2 launch computes the step's dt
// (every thread the same, driver_dt) and publishes it in st->dt; the launch
// with the epilogue ends the step: dt history, t += dt, CFL minimum reset.
// S: the storage type of the state and right-hand side (float in the FP32
// form); the update itself runs in double.
// The phase-clock build's record: per work-list entry of the update
// launches that run with the recording switched on (the last one wins), the
// clock readings relative to PC_START and the counts, for the first
// kPhaseCells entries.  The two lanes of a pair run together, so their
// cycles are the pair's; the settling count is the larger of the two lanes'
// fixed points.
#ifdef IMEX_PHASE_CLOCKS
constexpr int kPhaseCells = 65536;
__device__ unsigned long long g_phase[kPhaseCells * PC_NSLOT];
__device__ int g_phase_on;
__device__ __forceinline__ void record_phase_clocks(const PhaseClk& pc, int i, int lane) {
  if (!g_phase_on || i >= kPhaseCells) return;
  unsigned long long* r = g_phase + (size_t)i * PC_NSLOT;
  const unsigned long long s_its = pc.t[PC_SETTLE_ITS];
  const unsigned long long s_other = __shfl_xor(s_its, 1);
  if (lane != 0) return;
  for (int k = 1; k <= PC_END; k++) r[k] = pc.t[k] ? pc.t[k] - pc.t[PC_START] : 0ULL;
  r[PC_START] = 0ULL;
  for (int k = PC_ITS; k < PC_NSLOT; k++) r[k] = pc.t[k];
  r[PC_SETTLE_ITS] = (s_its > s_other) ? s_its : s_other;  // the pair's longer fixed point
}
#else
__device__ __forceinline__ void record_phase_clocks(const PhaseClk&, int, int) {}
#endif

template <int NS, int PARTS, typename S = double>
__device__ __forceinline__ void update_body(const Params& P, const Grid& G, double dt_arg, int mode,
                                            int step_idx) {
  constexpr int NV = 4 + NS;
  constexpr bool ALL = (PARTS & 8) != 0;
  // lanes per cell: the stage-2 launch evaluates the two Jacobian columns of
  // the Newton solve on an aligned lane pair, the sedimentation launch splits
  // the solid classes' settling velocities over a pair; the pair's lead lane
  // writes the results
  constexpr int LANES = (!ALL && ((PARTS & 1) || ((PARTS & 2) && NS > 1))) ? 2 : 1;
  const int lane = (LANES == 2) ? (int)(threadIdx.x & 1u) : 0;
  const bool lead = lane == 0;
  const int N = G.N;
  double dt = dt_arg;
  if (mode == 1) {
    if constexpr ((PARTS & 1) != 0) {
      if (G.st->done) {  // t reached t_end at an earlier step of the call
        if (blockIdx.x == 0 && threadIdx.x == 0 && G.dts) G.dts[step_idx] = 0.0;
        return;
      }
      dt = driver_dt(P, G.st);
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        G.st->stop_next = (G.st->t + dt < G.st->t_end) ? 0u : 1u;
        G.st->dt = dt;
        G.st->dt_cfl = fmin_f(P.max_dt, __longlong_as_double((long long)G.st->dtmin_bits));
        if (G.dts) G.dts[step_idx] = dt;
        // the whole step in this launch: other blocks may still be reading
        // t and the dt history, so the commit waits for the next launch
        if constexpr ((PARTS & 4) != 0) G.st->pending = 1u;
      }
    } else {
      dt = G.st->dt;
    }
  }
  const int n = ALL ? N : (int)G.st->n_work;
  if constexpr ((PARTS & 1) != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) G.st->work += (unsigned long long)n;
  }
  unsigned int err = 0;
  int dried = 0;
  const int stride = (int)(gridDim.x * NTHR) / LANES;
  // the x**y tables of the sedimentation in LDS, in blocks that have cells
  // (the condition is the same for every thread of a block)
  __shared__ ddm::dd s_pow[((PARTS & 2) != 0) ? ddm::POW_TAB_N : 1];
  if constexpr ((PARTS & 2) != 0) {
    if ((int)(blockIdx.x * NTHR) / LANES < n) pow_tables_to_lds(s_pow);
  }
  constexpr bool SED = (PARTS & 2) != 0;
  const bool sed = SED && (P.settling || P.entrainment || P.erosion_coeff > 0.0);
  const bool er = sed && G.erodible != nullptr;
  for (int i = (int)(blockIdx.x * NTHR + threadIdx.x) / LANES; i < n; i += stride) {
    PhaseClk pcl{};
    PhaseClk* pc = kPhaseClocks ? &pcl : nullptr;
    pc_mark(pc, PC_START);
    const int c = ALL ? i : G.work[i];
    double q[NV], X[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) q[v] = (double)state_q<S>(G)[(size_t)v * N + c];
    if constexpr ((PARTS & 1) != 0) {
#pragma unroll
      for (int v = 0; v < NV; v++) X[v] = (double)state_R<S>(G)[(size_t)v * N + c];
    }
    const bool wet0 = q[0] > 0.0;
    if constexpr ((PARTS & 1) != 0) {
      stage2<NS, LANES>(P, cell_xi(P, G.noise, c), dt, q, X, err, lane, pc);
      pc_mark(pc, PC_STAGE2END);
    }
    // The loads the rest of the cell's update needs are issued here, before
    // the sedimentation (whose settling fixed point covers their latency),
    // not where their values are first used: a load issued at its use would
    // wait for its memory latency on the launch's critical path.  They are
    // not issued before the Newton solve, whose registers are the kernel's
    // peak.
    double depo[NS], erod[NS], ero[NS], bc = 0.0;
    unsigned char m0 = 0;
    if constexpr (SED) {
#pragma unroll
      for (int v = 0; v < NS; v++) {
        depo[v] = sed ? G.dep[(size_t)v * N + c] : 0.0;
        erod[v] = er ? G.erodible[(size_t)v * N + c] : 0.0;
        ero[v] = er ? G.erosion[(size_t)v * N + c] : 0.0;
      }
      if (sed && P.topo_change) bc = G.bcent[c];
      // without the stage-2 part in this launch, "wet at the step start" is
      // bit 1 of the mask byte that launch left
      if constexpr ((PARTS & 1) == 0) m0 = G.mask[c];
    }
    double hm = 0.0, pm = 0.0, mv = 0.0;
    unsigned long long vb = 0ULL;
    if constexpr ((PARTS & 12) != 0) {
      hm = G.hmax[c];
      pm = G.pdyn[c];
      mv = G.mvel[c];
      vb = G.vuln ? G.vuln[c] : 0ULL;
    }
    if constexpr (SED) {
      if (sed) {
        double topo;
        pc_mark(pc, PC_SED0);
        erosion_deposition<NS, LANES>(P, dt, q, depo, erod, ero, topo, err, s_pow, lane, pc);
        pc_mark(pc, PC_SED1);
        if (P.topo_change) bc = bc + dt * topo;
      }
      // wet at the step start, exactly dry after the step's last change of q
      const bool was_wet = ((PARTS & 1) != 0) ? wet0 : ((m0 & 2u) != 0);
      if (lead && was_wet && q[0] == 0.0) dried++;
    }
    if constexpr ((PARTS & 12) != 0) {
      if (lead) epilogue<NS>(P, q, hm, pm, mv, vb, ALL);
    }
    // the stores, all after the arithmetic
    if (lead) {
      // "wet at the step start" in bit 1 of the mask byte, for the dried-cell
      // count of a sedimentation launch of its own
      if constexpr ((PARTS & 1) != 0) G.mask[c] = (unsigned char)(1u | (wet0 ? 2u : 0u));
      if constexpr (SED) {
        if (sed) {
#pragma unroll
          for (int v = 0; v < NS; v++) G.dep[(size_t)v * N + c] = depo[v];
          if (P.topo_change) G.bcent[c] = bc;
          if (er) {
#pragma unroll
            for (int v = 0; v < NS; v++) {
              G.erodible[(size_t)v * N + c] = erod[v];
              G.erosion[(size_t)v * N + c] = ero[v];
            }
          }
        }
      }
      if constexpr ((PARTS & 3) != 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) state_q<S>(G)[(size_t)v * N + c] = S(q[v]);
      }
      if constexpr ((PARTS & 12) != 0) {
        G.hmax[c] = hm;
        G.pdyn[c] = pm;
        G.mvel[c] = mv;
        if (G.vuln) G.vuln[c] = vb;
      }
    }
    if constexpr (kPhaseClocks) {
      pc_mark(pc, PC_END);
      record_phase_clocks(pcl, i, lane);
    }
  }
  if (err && lead) atomicOr(&G.st->err, err);
  if (dried) atomicAdd(&G.st->dried, (unsigned long long)dried);
  if ((PARTS & 4) && !(PARTS & 1) && mode == 1 && blockIdx.x == 0 && threadIdx.x == 0) {
    DevState* st = G.st;
    st->dt_old_old = st->dt_old;
    st->dt_old = dt;
    st->t = st->t + dt;
    st->dtmin_bits = 0x7ff0000000000000ULL;
  }
}

template <int NS, int PARTS, typename S>
__global__ __launch_bounds__(NTHR) void k_update(Params P, Grid G, double dt_arg, int mode,
                                                 int step_idx) {
  update_body<NS, PARTS, S>(P, G, dt_arg, mode, step_idx);
}

// ===========================================================================
// layout conversion and diagnostics
// ===========================================================================
template <typename S>
__global__ void k_aos_to_soa(const double* aos, S* soa, int nv, int N) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < N)
    for (int i = 0; i < nv; i++) soa[(size_t)i * N + c] = aos[(size_t)c * nv + i];
}
template <typename S>
__global__ void k_soa_to_aos(const S* soa, double* aos, int nv, int N) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < N)
    for (int i = 0; i < nv; i++) aos[(size_t)c * nv + i] = soa[(size_t)i * N + c];
}

// Neumaier-compensated partial sums of the stdout diagnostics: each thread
// accumulates its cells as a (sum, compensation) pair, the block combines
// its threads' pairs in thread order, and the host combines the blocks'
// pairs; one pair per block and quantity.
__device__ __forceinline__ void neumaier_add(double& sum, double& comp, double x) {
  const double t = sum + x;
  if (fabs(sum) >= fabs(x))
    comp += (sum - t) + x;
  else
    comp += (x - t) + sum;
  sum = t;
}
template <int NS, typename S>
__global__ __launch_bounds__(256) void k_diag(Params P, Grid G, double* part) {
  constexpr int NV = 4 + NS, NP = NV + 2;
  __shared__ double sh[4][256], shc[4][256];
  const int tid = threadIdx.x;
  double acc[4] = {0, 0, 0, 0}, cmp[4] = {0, 0, 0, 0};
  for (int c = blockIdx.x * blockDim.x + tid; c < G.N; c += gridDim.x * blockDim.x) {
    double q[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = (double)state_q<S>(G)[(size_t)i * G.N + c];
    double h = 0.0;
    if (q[0] > 0.0) {
      double qp[NP], pd;
      qc_to_qp<NS>(P, q, qp, pd);
      h = qp[0];
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) s += q[4 + i];
    neumaier_add(acc[0], cmp[0], q[0]);
    neumaier_add(acc[1], cmp[1], h);
    neumaier_add(acc[2], cmp[2], (q[0] > 1.0e-7) ? 1.0 : 0.0);
    neumaier_add(acc[3], cmp[3], s);
  }
  for (int m = 0; m < 4; m++) {
    sh[m][tid] = acc[m];
    shc[m][tid] = cmp[m];
  }
  __syncthreads();
  if (tid < 4) {
    double sum = 0.0, comp = 0.0;
    for (int i = 0; i < 256; i++) {
      neumaier_add(sum, comp, sh[tid][i]);
      comp += shc[tid][i];
    }
    part[(blockIdx.x * 4 + tid) * 2] = sum;
    part[(blockIdx.x * 4 + tid) * 2 + 1] = comp;
  }
}

// The host's combination of k_diag's per-block pairs (imex_post_step), on
// the device: thread m sums quantity m over the blocks in block order with
// the same Neumaier steps; out[m] = sum + compensation.
__global__ void k_diag_finish(const double* part, int nb, double* out) {
  const int m = threadIdx.x;
  if (m >= 4) return;
  double sum = 0.0, comp = 0.0;
  for (int b = 0; b < nb; b++) {
    const double x = part[(b * 4 + m) * 2];
    comp += part[(b * 4 + m) * 2 + 1];
    const double t = sum + x;
    if (fabs(sum) >= fabs(x))
      comp += (sum - t) + x;
    else
      comp += (x - t) + sum;
    sum = t;
  }
  out[m] = sum + comp;
}

// ===========================================================================
// per-function batch kernels (known-answer tests)
// ===========================================================================
template <int NS>
__global__ void kt_qc_to_qp(Params P, int n, const double* qc, double* qp, double* pd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) qc_to_qp<NS>(P, qc + (size_t)(4 + NS) * i, qp + (size_t)(6 + NS) * i, pd[i]);
}
template <int NS>
__global__ void kt_qp_to_qc(Params P, int n, const double* qp, double* qc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) qp_to_qc<NS>(P, qp + (size_t)(6 + NS) * i, qc + (size_t)(4 + NS) * i);
}
template <int NS>
__global__ void kt_fluxes(Params P, int n, int dir, const double* qc, const double* qp, double* f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) phys_flux<NS>(P, qc + (size_t)(4 + NS) * i, qp + (size_t)(6 + NS) * i, dir, f + (size_t)(4 + NS) * i);
}
template <int NS>
__global__ void kt_speeds(Params P, int n, int dir, const double* qp, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) local_speeds<NS>(P, qp + (size_t)(6 + NS) * i, dir, out[2 * i], out[2 * i + 1]);
}
template <int NS>
__global__ void kt_implicit_cplx(Params P, int n, const double* qre, const double* qim, double* ore,
                                 double* oim) {
  constexpr int NV = 4 + NS;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    cx q[NV];
    for (int m = 0; m < NV; m++) q[m] = C(qre[NV * i + m], qim[NV * i + m]);
    cx n1, n2;
    implicit_c<NS>(P, q, n1, n2, P.xi);
    for (int m = 0; m < NV; m++) {
      ore[NV * i + m] = (m == 1) ? n1.re : ((m == 2) ? n2.re : 0.0);
      oim[NV * i + m] = (m == 1) ? n1.im : ((m == 2) ? n2.im : 0.0);
    }
  }
}
// the residual's implicit terms as eval_f computes them (real path, complex
// fallback); fast[i] = 1 when the real path was taken
template <int NS>
__global__ void kt_implicit_real(Params P, int n, const double* q, double* out, int* fast) {
  constexpr int NV = 4 + NS;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double n1, n2;
    const bool ok = implicit_r<NS>(P, q + (size_t)NV * i, n1, n2, P.xi);
    if (!ok) {
      cx qc[NV];
      for (int m = 0; m < NV; m++) qc[m] = C(q[(size_t)NV * i + m], 0.0);
      cx c1, c2;
      implicit_c<NS>(P, qc, c1, c2, P.xi);
      n1 = c1.re;
      n2 = c2.re;
    }
    for (int m = 0; m < NV; m++) out[(size_t)NV * i + m] = (m == 1) ? n1 : ((m == 2) ? n2 : 0.0);
    fast[i] = ok ? 1 : 0;
  }
}
template <int NS>
__global__ void kt_semi_impl(Params P, int n, const double* qp, double* si) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) semi_impl<NS>(P, qp + (size_t)(6 + NS) * i, si + (size_t)(4 + NS) * i);
}
template <int NS>
__global__ void kt_expl(Params P, int n, const double* bx, const double* by, double t,
                        const double* qp, double* ex) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) expl_terms<NS>(P, bx[i], by[i], qp + (size_t)(6 + NS) * i, ex + (size_t)(4 + NS) * i, t, 0.0, 0.0, 0.0);
}
__global__ void kt_settling(Params P, int n, const double* d, const double* rs, const double* rc,
                            const double* ikv, double* out) {
  __shared__ ddm::dd tab[ddm::POW_TAB_N];
  pow_tables_to_lds(tab);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = settling_velocity(P, d[i], rs[i], rc[i], ikv[i], tab);
}
template <int NS>
__global__ void kt_mass_exchange(Params P, int n, const double* qp, const double* erod, double dt,
                                 double* dep, double* eq, double* ers) {
  __shared__ ddm::dd tab[ddm::POW_TAB_N];
  pow_tables_to_lds(tab);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double e[NS], r[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) e[k] = erod ? erod[(size_t)NS * i + k] : 0.0;
    double topo;
    mass_exchange<NS>(P, qp + (size_t)(6 + NS) * i, e, dt, dep + (size_t)NS * i,
                      eq + (size_t)(4 + NS) * i, r, topo, tab);
    if (ers)
#pragma unroll
      for (int k = 0; k < NS; k++) ers[(size_t)NS * i + k] = r[k];
  }
}
template <int NS>
__global__ void kt_newton(Params P, int n, double dt, const double* qsi, const double* q0,
                          const double* R, double* q, int* its) {
  constexpr int NV = 4 + NS;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double qq[NV];
    for (int m = 0; m < NV; m++) qq[m] = qsi[NV * i + m];
    unsigned int err = 0;
    its[i] = solve_rk_step<NS, 1>(P, P.xi, dt, qq, q0 + (size_t)NV * i, R + (size_t)NV * i, err, 0);
    for (int m = 0; m < NV; m++) q[NV * i + m] = qq[m];
  }
}
__global__ void kt_arith(int n, const double* a, const double* b, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    out[4 * i] = a[i] / b[i];
    out[4 * i + 1] = sqrt(fabs(a[i]));
    cx z = csqrt_c(C(a[i], b[i]));
    out[4 * i + 2] = z.re;
    out[4 * i + 3] = z.im;
  }
}

// even elements through the tables in LDS, odd ones through those in global
// memory: both lookups the kernels use
__global__ void kt_pow(int n, const double* x, const double* y, double* out) {
  __shared__ ddm::dd tab[ddm::POW_TAB_N];
  pow_tables_to_lds(tab);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (i & 1) ? pow_cr(x[i], y[i]) : pow_cr(x[i], y[i], tab);
}

// ===========================================================================
// host side
// ===========================================================================
#include "imex_handle.h"

// Points the kernels' per-step dt output at the handle's device buffer (grown
// to nsteps when needed) when the caller wants the dt sequence, else at NULL.
static int bind_dts(imex_handle* h, const double* dts, int nsteps) {
  if (dts && nsteps > h->dts_cap) {
    double* p = nullptr;
    if (int rc = dalloc(h, &p, (size_t)nsteps)) return rc;
    if (h->dts_buf) {
      HCHK(h, hipStreamSynchronize(h->stream));
      if (int rc = dfree(h, h->dts_buf)) return rc;
    }
    h->dts_buf = p;
    h->dts_cap = nsteps;
  }
  h->G.dts = dts ? h->dts_buf : nullptr;
  return 0;
}

template <int NS, typename F, int THR>
static size_t flux_lds_t() {
  if (THR == NTHR) return flux_lds_bytes<NS, F, NTHR>();
  // one one-pass set, or two NTHR-thread sets side by side
  return std::max(flux_lds_bytes<NS, F, 2 * NTHR>(), 2 * flux_lds_bytes<NS, F, NTHR>());
}
template <int NS, int THR>
static size_t flux_lds(bool fp32) {
  return fp32 ? flux_lds_t<NS, float, THR>() : flux_lds_t<NS, double, THR>();
}

// dynamic LDS of k_flux (FLUX_THR threads) or, with batch set, of kb_flux
static size_t lds_for(int ns, bool fp32 = false, bool batch = false) {
  if (batch) {
    switch (ns) {
      case 1: return flux_lds<1, NTHR>(fp32);
      case 2: return flux_lds<2, NTHR>(fp32);
      default: return flux_lds<3, NTHR>(fp32);
    }
  }
  switch (ns) {
    case 1: return flux_lds<1, FLUX_THR>(fp32);
    case 2: return flux_lds<2, FLUX_THR>(fp32);
    default: return flux_lds<3, FLUX_THR>(fp32);
  }
}

// Active cells of each tile when no cell is wet: the frame and the interior
// cells within Manhattan distance 2 of it (the closest frame cell is straight
// along an axis), minus source-disc cells.  k_mask adds the total once per
// pass and each visited tile its cells beyond these.
static int frame_band(imex_handle* h, const signed char* scell) {
  const Grid& G = h->G;
  std::vector<int> cnt((size_t)h->ntiles, 0);
  unsigned long long total = 0;
  for (int k = 0; k < G.ny; k++)
    for (int j = 0; j < G.nx; j++) {
      const bool band = j <= 2 || j >= G.nx - 3 || k <= 2 || k >= G.ny - 3;
      if (!band) continue;
      if (h->P.radial && scell && scell[j + (size_t)G.nx * k] == 1) continue;
      cnt[(size_t)(j / TX) + (size_t)G.ntx * (k / TY)]++;
      total++;
    }
  HCHK(h, hipMemcpyAsync(h->G.tile_static, cnt.data(), 4 * cnt.size(), hipMemcpyHostToDevice,
                         h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  h->G.static_active = total;
  return 0;
}

// Device scratch of imex_post_step_noalloc: the diagnostics kernel's
// per-block (sum, compensation) pairs of the four quantities (the same
// layout as the handle's own partials buffer) and the four finished sums.
// One copy per GPU, held until the process ends; host code uses the address
// hipGetSymbolAddress gives on each device, never the array's name.
constexpr int kNoallocDiagBlocks = 512;
static __device__ __align__(256) double scratch[kNoallocDiagBlocks * 8 + 8];
constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

// the address of `scratch` on the current device h->device (looked up once)
static int scratch_address(imex_handle* h, double** ptr) {
  if (h->device < 0 || h->device >= kMaxDevices) {
    h->err = "device index beyond the scratch table";
    return -2;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[h->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return set_err(h, "hipGetSymbolAddress(scratch)", e);
    g_scratch_dev[h->device] = static_cast<double*>(p);
  }
  *ptr = g_scratch_dev[h->device];
  return 0;
}

extern "C" int imex_create(const imex_config* cfg, imex_handle** out) {
  if (!cfg || !out) return -2;
  *out = nullptr;
  if (cfg->nx < 5 || cfg->ny < 5 || cfg->n_solid < 1 || cfg->n_solid > 3 ||
      cfg->n_thickness_levels < 0 || cfg->n_thickness_levels > 8 || cfg->n_dyn_pres_levels < 0 ||
      cfg->n_dyn_pres_levels > 8 || cfg->precision < 0 || cfg->precision > 1)
    return -3;
  // friction law: Voellmy-Salm (0 = 1) or mu(I) with non-negative parameters
  // (src/inpout_2d.f90:3592-3630)
  const int rheo = cfg->rheology_model == 0 ? 1 : cfg->rheology_model;
  if (rheo != 1 && rheo != 11 && rheo != 12) return -3;
  if (rheo != 1 && !(cfg->mu_s >= 0.0 && cfg->mu_2 >= 0.0 && cfg->I_0 >= 0.0 && cfg->muI_inf >= 0.0))
    return -3;
  // pore pressure (src/inpout_2d.f90:681-710, :1989-2037): the general path
  // of imex_pore.hip, device-driven steps, without the features it does not
  // carry
  if (cfg->pore_pressure_flag) {
    if (cfg->f_inhibit_mode < 0 || cfg->f_inhibit_mode > 3 || cfg->radial_source_flag ||
        cfg->settling_flag || cfg->entrainment_flag || cfg->erosion_coeff != 0.0 ||
        cfg->topo_change_flag || cfg->precision != 0)
      return -3;
    const double mp = cfg->maximum_solid_packing > 0.0 ? cfg->maximum_solid_packing : 0.64;
    if (cfg->f_inhibit_mode >= 2 &&
        (cfg->N_inh <= 0 || cfg->N_inh > 15 || !(cfg->alpha_trans > 0.0) || mp < cfg->alpha_trans))
      return -3;
  }
  // the transported noise (src/inpout_2d.f90:1466-1475): with the
  // stochastic friction, on the general path, which takes what the
  // pore-pressure runs take; the boundary tables hold 8 variables
  if (cfg->stoch_transport_flag) {
    if (!cfg->stochastic_flag || cfg->radial_source_flag || cfg->settling_flag ||
        cfg->entrainment_flag || cfg->erosion_coeff != 0.0 || cfg->topo_change_flag ||
        cfg->precision != 0 || 4 + cfg->n_solid + 1 + (cfg->pore_pressure_flag ? 1 : 0) > 8)
      return -3;
  }
  // boundaries other than zero gradient: the general path only
  if (cfg->bc_custom) {
    if (!cfg->pore_pressure_flag && !cfg->stoch_transport_flag) return -3;
    for (int i = 0; i < 32; i++)
      if (cfg->bc_flag[i] < 0 || cfg->bc_flag[i] > 2) return -3;
  }
  imex_handle* h = new imex_handle();
  h->cfg = *cfg;
  h->fp32 = cfg->precision == 1;
  const char* dev = getenv("IMEX_DEVICE");
  h->device = dev ? atoi(dev) : 0;
  {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) {
      set_err(h, "hipSetDevice", e);
      *out = h;
      return -1;
    }
    e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) {
      set_err(h, "hipStreamCreate", e);
      *out = h;
      return -1;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess &&
        ncu > 0)
      h->n_cu = ncu;
    double* sp = nullptr;
    if (scratch_address(h, &sp)) {
      *out = h;
      return -1;
    }
  }
  Params& P = h->P;
  memset(&P, 0, sizeof(P));
  P.nx = cfg->nx;
  P.ny = cfg->ny;
  P.ns = cfg->n_solid;
  P.nv = 4 + cfg->n_solid + (cfg->stoch_transport_flag ? 1 : 0) + (cfg->pore_pressure_flag ? 1 : 0);
  P.pore = cfg->pore_pressure_flag != 0;
  P.stoch_tr = cfg->stoch_transport_flag != 0;
  P.general = P.pore || P.stoch_tr;
  if (P.pore) {
    P.gas_loss = cfg->gas_loss_flag != 0;
    P.finh = cfg->f_inhibit_mode;
    P.N_inh = cfg->N_inh;
    P.dyn_perm = cfg->dynamic_permeability_flag != 0;
    P.hyd_perm = cfg->hydraulic_permeability;
    P.alpha_trans = cfg->alpha_trans;
    for (int i = 0; i < 3; i++) P.sph_s[i] = cfg->sphericity_s[i] > 0.0 ? cfg->sphericity_s[i] : 1.0;
    // precompute_pascal_coefficient (src/constitutive_2d.f90:419-466)
    for (int n = 0; n <= P.N_inh && n < 16; n++) {
      const int Nn = P.N_inh;
      double p1 = 1.0, p2 = 1.0;
      for (int k = 0; k < n; k++) p1 = p1 * (-(double)Nn - 1.0 - (double)k) / (double)(k + 1);
      for (int k = 0; k < Nn - n; k++) p2 = p2 * (2.0 * (double)Nn + 1.0 - (double)k) / (double)(k + 1);
      P.pascal[n] = p1 * p2;
    }
  }
  P.max_pack = cfg->maximum_solid_packing > 0.0 ? cfg->maximum_solid_packing : 0.64;
  for (int sd = 0; sd < 4; sd++)
    for (int i = 0; i < 8; i++) {
      P.bcf[sd][i] = cfg->bc_custom ? cfg->bc_flag[8 * sd + i] : 1;
      P.bcv[sd][i] = cfg->bc_custom ? cfg->bc_value[8 * sd + i] : 0.0;
    }
  P.energy = cfg->energy_flag;
  P.settling = cfg->settling_flag;
  P.entrainment = cfg->entrainment_flag;
  P.radial = cfg->radial_source_flag;
  P.n_thk = cfg->n_thickness_levels;
  P.n_pdyn = cfg->n_dyn_pres_levels;
  if (P.n_thk * P.n_pdyn > 64) P.n_pdyn = 64 / (P.n_thk ? P.n_thk : 1);
  P.dx = cfg->dx;
  P.dy = cfg->dx;
  P.dx2 = P.dx / 2.0;
  P.dy2 = P.dy / 2.0;
  P.one_by_dx = 1.0 / P.dx;
  P.one_by_dy = 1.0 / P.dy;
  P.cfl = cfg->cfl;
  P.max_dt = cfg->max_dt;
  P.grav = cfg->grav;
  P.mu = cfg->mu;
  P.xi = cfg->xi;
  P.pres = cfg->pres;
  P.inv_pres = 1.0 / cfg->pres;
  P.T_amb = cfg->T_ambient;
  P.cp_a = cfg->sp_heat_a;
  P.R_a = cfg->sp_gas_const_a;
  P.kin_visc_c = cfg->kin_visc_a;
  P.rho_a_amb = cfg->pres / (cfg->sp_gas_const_a * cfg->T_ambient);
  {
    const double d4 = pow(cfg->dx, 4.0);
    P.eps_sing = std::min(std::min(d4, d4), 1.0e-10);
    const double e = P.eps_sing;
    P.eps_sing4 = e * (e * (e * e));
  }
  for (int i = 0; i < 3; i++) {
    P.rho_s[i] = cfg->rho_s[i];
    P.inv_rho_s[i] = (i < cfg->n_solid) ? 1.0 / cfg->rho_s[i] : 0.0;
    P.diam_s[i] = cfg->diam_s[i];
    P.cp_s[i] = cfg->sp_heat_s[i];
    P.alphas_source[i] = cfg->alphas_source[i];
  }
  // erosion parameters as read_param leaves them (src/inpout_2d.f90:1827-1900)
  P.erosion_coeff = cfg->erosion_coeff;
  if (P.erosion_coeff > 0.0) {
    P.erodible_porosity = cfg->erodible_porosity;
    P.coeff_porosity = P.erodible_porosity / (1.0 - P.erodible_porosity);
    for (int i = 0; i < 3; i++) P.erodible_fract[i] = (cfg->n_solid == 1) ? (i == 0 ? 1.0 : 0.0)
                                                                           : cfg->erodible_fract[i];
    P.T_erodible = cfg->T_erodible;
    P.erodible_deposit = cfg->erodible_deposit_flag;
  } else {
    P.erodible_porosity = 0.0;
    P.coeff_porosity = 0.0;
    for (int i = 0; i < 3; i++) P.erodible_fract[i] = 1.0 / cfg->n_solid;
    P.T_erodible = 300.0;
    P.erodible_deposit = cfg->erodible_deposit_flag;
  }
  P.topo_change = cfg->topo_change_flag;
  P.rheology = rheo;
  P.stochastic = cfg->stochastic_flag != 0;
  P.mu_s = cfg->mu_s;
  P.mu_2 = cfg->mu_2;
  P.muI_inf = cfg->muI_inf;
  P.I_0 = cfg->I_0;
  P.rho_c_sub = cfg->pres / (cfg->sp_gas_const_a * P.T_erodible);
  P.vel_source = cfg->vel_source;
  P.T_source = cfg->T_source;
  P.h_source = cfg->h_source;
  for (int i = 0; i < 4; i++) P.time_param[i] = cfg->time_param[i];
  P.pi_g = 4.0 * atan(1.0);
  for (int i = 0; i < 8; i++) {
    P.thk[i] = cfg->thickness_levels[i];
    P.pdyn[i] = cfg->dyn_pres_levels[i];
  }
  P.hcs = P.nv * 2.220446049250313e-16;  // n_vars * EPSILON (src/solver_2d.f90:274-275)
  P.one_by_hcs = 1.0 / P.hcs;

  Grid& G = h->G;
  memset(&G, 0, sizeof(G));
  G.nx = cfg->nx;
  G.ny = cfg->ny;
  G.N = cfg->nx * cfg->ny;
  G.ntx = (cfg->nx + TX - 1) / TX;
  G.nty = (cfg->ny + TY - 1) / TY;
  h->ntiles = G.ntx * G.nty;
  const size_t N = (size_t)G.N;
  const int nv = P.nv, ns = P.ns;
  int rc = 0;
  *out = h;
  if (h->fp32) {
    if ((rc = dalloc(h, &G.qf, N * nv))) return rc;
    if ((rc = dalloc(h, &G.Rf, N * nv))) return rc;
  } else {
    if ((rc = dalloc(h, &G.q, N * nv))) return rc;
    if ((rc = dalloc(h, &G.R, N * nv))) return rc;
  }
  if ((rc = dalloc(h, &G.hmax, N))) return rc;
  if ((rc = dalloc(h, &G.pdyn, N))) return rc;
  if ((rc = dalloc(h, &G.mvel, N))) return rc;
  if ((rc = dalloc(h, &G.dep, N * ns))) return rc;
  if (P.erosion_coeff > 0.0 || P.erodible_deposit) {
    if ((rc = dalloc(h, &G.erodible, N * ns))) return rc;
    if ((rc = dalloc(h, &G.erosion, N * ns))) return rc;
  }
  if (P.topo_change)
    if ((rc = dalloc(h, &G.bcent, N))) return rc;
  if (P.stochastic)  // Z = 0 until imex_set_noise (the reference's allocation value)
    if ((rc = dalloc(h, &G.noise, N))) return rc;
  if (P.n_thk * P.n_pdyn > 0)
    if ((rc = dalloc(h, &G.vuln, N))) return rc;
  if ((rc = dalloc(h, &G.mask, N))) return rc;
  if ((rc = dalloc(h, &G.tile_list, 4 * (size_t)h->ntiles))) return rc;  // [2][2 ntiles] half tiles
  if ((rc = dalloc(h, &G.work, N))) return rc;
  if ((rc = dalloc(h, &G.cand, 3 * (size_t)h->ntiles))) return rc;
  if ((rc = dalloc(h, &G.cand_stamp, (size_t)h->ntiles))) return rc;
  if ((rc = dalloc(h, &G.tile_static, (size_t)h->ntiles))) return rc;
  double *bx, *by, *ap, *ax, *ay, *xc, *xs, *yc, *ys;
  signed char* sc;
  unsigned char* sd;
  if ((rc = dalloc(h, &bx, N))) return rc;
  if ((rc = dalloc(h, &by, N))) return rc;
  if ((rc = dalloc(h, &ap, N))) return rc;
  if ((rc = dalloc(h, &ax, N))) return rc;
  if ((rc = dalloc(h, &ay, N))) return rc;
  if ((rc = dalloc(h, &sc, N))) return rc;
  if ((rc = dalloc(h, &sd, N))) return rc;
  if ((rc = dalloc(h, &xc, (size_t)G.nx))) return rc;
  if ((rc = dalloc(h, &xs, (size_t)G.nx + 1))) return rc;
  if ((rc = dalloc(h, &yc, (size_t)G.ny))) return rc;
  if ((rc = dalloc(h, &ys, (size_t)G.ny + 1))) return rc;
  G.bx = bx; G.by = by; G.arcp = ap; G.arcx = ax; G.arcy = ay;
  G.scell = sc; G.sides = sd; G.xc = xc; G.xs = xs; G.yc = yc; G.ys = ys;
  DevState* st;
  if ((rc = dalloc(h, &st, 1))) return rc;
  G.st = st;
  if ((rc = dalloc(h, &h->d_tmp, N * nv))) return rc;
  h->diag_blocks = 512;
  if ((rc = dalloc(h, &h->d_diag, (size_t)h->diag_blocks * 8))) return rc;
  // grid coordinates as init_grid builds them (src/geometry_2d.f90:312-330)
  {
    std::vector<double> hxs(G.nx + 1), hxc(G.nx), hys(G.ny + 1), hyc(G.ny);
    for (int j = 0; j <= G.nx; j++) hxs[j] = cfg->x0 + j * cfg->dx;
    for (int k = 0; k <= G.ny; k++) hys[k] = cfg->y0 + k * cfg->dx;
    for (int j = 0; j < G.nx; j++) hxc[j] = 0.5 * (hxs[j] + hxs[j + 1]);
    for (int k = 0; k < G.ny; k++) hyc[k] = 0.5 * (hys[k] + hys[k + 1]);
    HCHK(h, hipMemcpyAsync(xs, hxs.data(), 8 * hxs.size(), hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(xc, hxc.data(), 8 * hxc.size(), hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(ys, hys.data(), 8 * hys.size(), hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipMemcpyAsync(yc, hyc.data(), 8 * hyc.size(), hipMemcpyHostToDevice, h->stream));
    DevState s0;
    memset(&s0, 0, sizeof(s0));
    s0.dtmin_bits = 0x7ff0000000000000ULL;
    HCHK(h, hipMemcpyAsync(st, &s0, sizeof(s0), hipMemcpyHostToDevice, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
  }
  if ((rc = frame_band(h, nullptr))) return rc;
  if (P.general && (rc = imex_pore_alloc(h))) return rc;
  h->lds = lds_for(ns, h->fp32);
  {
    int max_lds = 0;
    HCHK(h, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    if ((size_t)max_lds < h->lds) {
      h->err = "k_flux needs " + std::to_string(h->lds) + " bytes of LDS per block, the device allows " +
               std::to_string(max_lds);
      *out = h;
      return -1;
    }
  }
  switch (ns) {
    case 1: HCHK(h, hipFuncSetAttribute(h->fp32 ? (const void*)k_flux<1, float> : (const void*)k_flux<1, double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds)); break;
    case 2: HCHK(h, hipFuncSetAttribute(h->fp32 ? (const void*)k_flux<2, float> : (const void*)k_flux<2, double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds)); break;
    default: HCHK(h, hipFuncSetAttribute(h->fp32 ? (const void*)k_flux<3, float> : (const void*)k_flux<3, double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds)); break;
  }
  return 0;
}

extern "C" int imex_set_static(imex_handle* h, const imex_static* s) {
  if (!h || !s || !s->B_prime_x || !s->B_prime_y) return -2;
  HCHK(h, hipSetDevice(h->device));
  const size_t N = (size_t)h->G.N;
  std::vector<signed char> sc(N, 0);
  std::vector<unsigned char> sd(N, 0);
  for (size_t c = 0; c < N; c++) {
    if (s->source_cell) sc[c] = (signed char)s->source_cell[c];
    if (s->source_sides) sd[c] = (unsigned char)s->source_sides[c];
  }
  HCHK(h, hipMemcpyAsync((void*)h->G.bx, s->B_prime_x, 8 * N, hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipMemcpyAsync((void*)h->G.by, s->B_prime_y, 8 * N, hipMemcpyHostToDevice, h->stream));
  if (s->arc_perim)
    HCHK(h, hipMemcpyAsync((void*)h->G.arcp, s->arc_perim, 8 * N, hipMemcpyHostToDevice, h->stream));
  if (s->arc_n_x)
    HCHK(h, hipMemcpyAsync((void*)h->G.arcx, s->arc_n_x, 8 * N, hipMemcpyHostToDevice, h->stream));
  if (s->arc_n_y)
    HCHK(h, hipMemcpyAsync((void*)h->G.arcy, s->arc_n_y, 8 * N, hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipMemcpyAsync((void*)h->G.scell, sc.data(), N, hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipMemcpyAsync((void*)h->G.sides, sd.data(), N, hipMemcpyHostToDevice, h->stream));
  if (int rc = frame_band(h, sc.data())) return rc;
  HCHK(h, hipStreamSynchronize(h->stream));
  h->mask_full_next = true;
  return 0;
}

// Persistent grids: enough blocks to fill the 256 CUs at the kernels'
// occupancy; each block walks the active-tile list with a stride of the grid.
static int grid_for(const imex_handle* h, int per_cu, int items = -1) {
  return std::max(1, std::min(items < 0 ? h->ntiles : items, h->n_cu * per_cu));
}

template <int NS, typename S>
static void launch_update(imex_handle* h, int parts, double dt, int mode, int step_idx) {
  const int nb = (parts & 8) ? (h->G.N + NTHR - 1) / NTHR : grid_for(h, 4);
  const dim3 g(nb), b(NTHR);
  const Params& P = h->P;
  const Grid& G = h->G;
  hipStream_t s = h->stream;
  switch (parts) {
    case 1: hipLaunchKernelGGL((k_update<NS, 1, S>), g, b, 0, s, P, G, dt, mode, step_idx); break;
    case 2: hipLaunchKernelGGL((k_update<NS, 2, S>), g, b, 0, s, P, G, dt, mode, step_idx); break;
    case 4: hipLaunchKernelGGL((k_update<NS, 4, S>), g, b, 0, s, P, G, dt, mode, step_idx); break;
    case 6: hipLaunchKernelGGL((k_update<NS, 6, S>), g, b, 0, s, P, G, dt, mode, step_idx); break;
    case 7: hipLaunchKernelGGL((k_update<NS, 7, S>), g, b, 0, s, P, G, dt, mode, step_idx); break;
    default: hipLaunchKernelGGL((k_update<NS, 8, S>), g, b, 0, s, P, G, dt, mode, step_idx); break;
  }
}
// parts: 1, 2, 4, 8 as the host-driven entry points call them, or 1|2|4 for
// a device-driven step (mode 1), which runs as two launches (1, then 2|4).
static int update(imex_handle* h, int parts, double dt, int mode, int step_idx = 0) {
  switch (h->P.ns) {
    case 1:
      if (h->fp32) launch_update<1, float>(h, parts, dt, mode, step_idx);
      else launch_update<1, double>(h, parts, dt, mode, step_idx);
      break;
    case 2:
      if (h->fp32) launch_update<2, float>(h, parts, dt, mode, step_idx);
      else launch_update<2, double>(h, parts, dt, mode, step_idx);
      break;
    default:
      if (h->fp32) launch_update<3, float>(h, parts, dt, mode, step_idx);
      else launch_update<3, double>(h, parts, dt, mode, step_idx);
      break;
  }
  HCHK(h, hipGetLastError());
  return 0;
}
// topography_reconstruction's bed slopes (src/geometry_2d.f90:744-836) from
// B_cent, every cell: the 5-point stencil (-2, -1, 0, 1, 2) / (10 dx) summed
// in the reference's order at interior cells, and the border copies, which
// give each cell the value of the nearest interior cell (x, then y).
__global__ void k_topo(Params P, Grid G) {
  const double c1[5] = {-2.0, -1.0, 0.0, 1.0, 2.0};
  const int nx = G.nx, ny = G.ny;
  const double nxn = 10.0 * P.dx, nyn = 10.0 * P.dy;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < G.N; c += gridDim.x * blockDim.x) {
    const int j = c % nx, k = c / nx;
    const int js = min(max(j, 2), nx - 3), ks = min(max(k, 2), ny - 3);
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int m = 0; m < 5; m++) sx = sx + c1[m] * G.bcent[(js + m - 2) + (size_t)nx * ks];
#pragma unroll
    for (int m = 0; m < 5; m++) sy = sy + c1[m] * G.bcent[js + (size_t)nx * (ks + m - 2)];
    G.bx[c] = sx / nxn;
    G.by[c] = sy / nyn;
  }
}
// a run with topography change steps only once B_cent is on the device
static int need_topography(imex_handle* h) {
  if (h->P.topo_change && !h->bcent_set) {
    h->err = "topo_change_flag: imex_set_topography must give B_cent before the first step";
    return -2;
  }
  return 0;
}
static int topo(imex_handle* h) {
  if (!h->P.topo_change) return 0;
  const int nb = std::max(1, std::min(8 * h->n_cu, (h->G.N + 255) / 256));
  hipLaunchKernelGGL(k_topo, dim3(nb), dim3(256), 0, h->stream, h->P, h->G);
  HCHK(h, hipGetLastError());
  return 0;
}

// applies a pending end of step (after the last device-driven step of a call)
static int commit(imex_handle* h) {
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(1), 0, h->stream, h->G);
  HCHK(h, hipGetLastError());
  return 0;
}
static int mask(imex_handle* h, int solve_all, int driven = 0) {
  h->par ^= 1;
  h->mask_seq++;
  const int full = (solve_all || h->mask_full_next) ? 1 : 0;
  h->mask_full_next = solve_all != 0;  // the pass after solve_all redoes every tile
  const int nb = full ? h->ntiles : grid_for(h, 4);
  hipLaunchKernelGGL((h->fp32 ? k_mask<float> : k_mask<double>), dim3(nb), dim3(NTHR), 0, h->stream, h->P, h->G, solve_all, h->par,
                     h->mask_seq, full, driven);
  HCHK(h, hipGetLastError());
  return 0;
}
static int flux(imex_handle* h, int mode, double t, int step_idx) {
  // one block of FLUX_THR threads per CU (its registers and LDS set fill the
  // CU): the blocks are all resident, so a half tile per block is one round
  const int nb = grid_for(h, 1, 2 * h->ntiles);  // work items are half tiles
  switch (h->P.ns) {
    case 1: hipLaunchKernelGGL((h->fp32 ? k_flux<1, float> : k_flux<1, double>), dim3(nb), dim3(FLUX_THR), h->lds, h->stream, h->P, h->G, mode, t, h->par); break;
    case 2: hipLaunchKernelGGL((h->fp32 ? k_flux<2, float> : k_flux<2, double>), dim3(nb), dim3(FLUX_THR), h->lds, h->stream, h->P, h->G, mode, t, h->par); break;
    default: hipLaunchKernelGGL((h->fp32 ? k_flux<3, float> : k_flux<3, double>), dim3(nb), dim3(FLUX_THR), h->lds, h->stream, h->P, h->G, mode, t, h->par); break;
  }
  HCHK(h, hipGetLastError());
  (void)step_idx;
  if (mode == 0) {  // the device-side driver takes its dt in k_update instead
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, h->stream, h->P, h->G);
    HCHK(h, hipGetLastError());
  }
  return 0;
}

static int read_state(imex_handle* h, DevState* s) {
  HCHK(h, hipMemcpyAsync(s, h->G.st, sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int imex_upload_state(imex_handle* h, const double* q_aos, double t, double dt0) {
  if (!h || !q_aos) return -2;
  HCHK(h, hipSetDevice(h->device));
  const int N = h->G.N, nv = h->P.nv;
  // a dry cell (q1 <= 0) with other non-zero components is zeroed by the
  // assembly the first time it is active; until then k_flux must not skip it
  h->P.all_tiles = 0;
  for (size_t c = 0; c < (size_t)N && !h->P.all_tiles; c++)
    if (!(q_aos[c * nv] > 0.0))
      for (int i = 1; i < nv; i++)
        if (q_aos[c * nv + i] != 0.0) {
          h->P.all_tiles = 1;
          break;
        }
  HCHK(h, hipMemcpyAsync(h->d_tmp, q_aos, 8 * (size_t)N * nv, hipMemcpyHostToDevice, h->stream));
  if (h->fp32)
    hipLaunchKernelGGL(k_aos_to_soa<float>, dim3((N + 255) / 256), dim3(256), 0, h->stream, h->d_tmp, h->G.qf, nv, N);
  else
    hipLaunchKernelGGL(k_aos_to_soa<double>, dim3((N + 255) / 256), dim3(256), 0, h->stream, h->d_tmp, h->G.q, nv, N);
  HCHK(h, hipGetLastError());
  DevState s;
  if (read_state(h, &s)) return -1;
  s.t = t;
  s.dt = dt0;
  s.dt_old = dt0;
  s.dt_old_old = dt0;
  s.dtmin_bits = 0x7ff0000000000000ULL;
  s.err = 0;
  s.pending = 0;
  s.done = 0;
  s.stop_next = 0;
  s.steps = 0;
  s.hot = 0;
  HCHK(h, hipMemcpyAsync(h->G.st, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
  int rc;
  if ((rc = mask(h, 1))) return rc;
  if ((rc = update(h, 8, 0.0, 0))) return rc;
  HCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// runs on the general path (pore pressure, the transported noise) step
// through imex_run_steps only
static int device_driven_only(imex_handle* h) {
  if (h->P.general) {
    h->err = "pore_pressure_flag / stoch_transport_flag: these runs take device-driven steps "
             "(imex_run_steps)";
    return -2;
  }
  return 0;
}

extern "C" int imex_check_solve(imex_handle* h, int solve_all) {
  if (!h) return -2;
  if (int rc = device_driven_only(h)) return rc;
  HCHK(h, hipSetDevice(h->device));
  return mask(h, solve_all);
}

extern "C" int imex_timestep(imex_handle* h, double t, double* dt_out) {
  if (!h) return -2;
  if (int rc = device_driven_only(h)) return rc;
  HCHK(h, hipSetDevice(h->device));
  int rc;
  if ((rc = flux(h, 0, t, 0))) return rc;
  DevState s;
  if (read_state(h, &s)) return -1;
  if (dt_out) *dt_out = s.dt_cfl;
  return 0;
}

static int take_err(imex_handle* h, unsigned* out) {
  DevState s;
  if (read_state(h, &s)) return -1;
  *out = s.err;
  return 0;
}

extern "C" int imex_rk_step(imex_handle* h, double dt) {
  if (!h) return -2;
  if (int rc = device_driven_only(h)) return rc;
  HCHK(h, hipSetDevice(h->device));
  return update(h, 1, dt, 0);
}

extern "C" int imex_erosion_deposition(imex_handle* h, double dt) {
  if (!h) return -2;
  if (int rc = device_driven_only(h)) return rc;
  if (int rc = need_topography(h)) return rc;
  HCHK(h, hipSetDevice(h->device));
  return update(h, 2, dt, 0);
}

extern "C" int imex_post_step(imex_handle* h, double t_new, imex_step_stats* out) {
  if (!h) return -2;
  HCHK(h, hipSetDevice(h->device));
  int rc;
  // runs on the general path (pore pressure, the transported noise) had
  // their epilogue inside each device-driven step; the diagnostics below
  // read only h, q1 and the solid masses, the same in their layout
  if (!h->P.general && (rc = update(h, 4, 0.0, 0))) return rc;
  DevState s;
  if (read_state(h, &s)) return -1;
  s.t = t_new;
  unsigned err = s.err;
  s.err = 0;
  HCHK(h, hipMemcpyAsync(h->G.st, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
  if (out) {
    switch (h->P.ns) {
      case 1: hipLaunchKernelGGL((h->fp32 ? k_diag<1, float> : k_diag<1, double>), dim3(h->diag_blocks), dim3(256), 0, h->stream, h->P, h->G, h->d_diag); break;
      case 2: hipLaunchKernelGGL((h->fp32 ? k_diag<2, float> : k_diag<2, double>), dim3(h->diag_blocks), dim3(256), 0, h->stream, h->P, h->G, h->d_diag); break;
      default: hipLaunchKernelGGL((h->fp32 ? k_diag<3, float> : k_diag<3, double>), dim3(h->diag_blocks), dim3(256), 0, h->stream, h->P, h->G, h->d_diag); break;
    }
    HCHK(h, hipGetLastError());
    std::vector<double> part((size_t)h->diag_blocks * 8);
    HCHK(h, hipMemcpyAsync(part.data(), h->d_diag, 8 * part.size(), hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    double tot[4];
    for (int m = 0; m < 4; m++) {
      double sum = 0.0, comp = 0.0;
      for (int b = 0; b < h->diag_blocks; b++) {
        const double x = part[(b * 4 + m) * 2];
        comp += part[(b * 4 + m) * 2 + 1];
        const double t = sum + x;
        if (fabs(sum) >= fabs(x))
          comp += (sum - t) + x;
        else
          comp += (x - t) + sum;
        sum = t;
      }
      tot[m] = sum + comp;
    }
    const double a = h->P.dx * h->P.dy;
    out->t = t_new;
    out->dt = s.dt;
    out->mass = a * tot[0];
    out->volume = a * tot[1];
    out->area = a * tot[2];
    out->solid_mass = a * tot[3];
    out->active_cells = (int64_t)s.active;
    out->dried_cells = (int64_t)s.dried;
    out->err = err;
  } else {
    HCHK(h, hipStreamSynchronize(h->stream));
  }
  return (int)(err & ~IMEX_WARN_RESCALE);
}

// imex_post_step without allocations: the diagnostics' partial sums and
// their combination live in the file-scope device array `scratch` (at most
// kNoallocDiagBlocks blocks of partials), finished on the device by
// k_diag_finish; the host reads back the driver state and four sums into
// stack memory.  Calls that overlap on one GPU share `scratch` and race.
extern "C" int imex_post_step_noalloc(imex_handle* h, double t_new, imex_step_stats* out) {
  if (!h) return -2;
  HCHK(h, hipSetDevice(h->device));
  double* sp = nullptr;
  if (scratch_address(h, &sp)) return -1;
  if (h->diag_blocks > kNoallocDiagBlocks) {
    h->err = "imex_post_step_noalloc: " + std::to_string(h->diag_blocks) +
                    " diagnostics blocks exceed the scratch of " +
                    std::to_string(kNoallocDiagBlocks);
    return -2;
  }
  int rc;
  // runs on the general path (pore pressure, the transported noise) had
  // their epilogue inside each device-driven step; the diagnostics below
  // read only h, q1 and the solid masses, the same in their layout
  if (!h->P.general && (rc = update(h, 4, 0.0, 0))) return rc;
  DevState s;
  if (read_state(h, &s)) return -1;
  s.t = t_new;
  const unsigned err = s.err;
  s.err = 0;
  HCHK(h, hipMemcpyAsync(h->G.st, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
  if (out) {
    const int nb = h->diag_blocks;
    switch (h->P.ns) {
      case 1: hipLaunchKernelGGL((h->fp32 ? k_diag<1, float> : k_diag<1, double>), dim3(nb), dim3(256), 0, h->stream, h->P, h->G, sp); break;
      case 2: hipLaunchKernelGGL((h->fp32 ? k_diag<2, float> : k_diag<2, double>), dim3(nb), dim3(256), 0, h->stream, h->P, h->G, sp); break;
      default: hipLaunchKernelGGL((h->fp32 ? k_diag<3, float> : k_diag<3, double>), dim3(nb), dim3(256), 0, h->stream, h->P, h->G, sp); break;
    }
    HCHK(h, hipGetLastError());
    double* fin = sp + (size_t)kNoallocDiagBlocks * 8;
    hipLaunchKernelGGL(k_diag_finish, dim3(1), dim3(64), 0, h->stream, sp, nb, fin);
    HCHK(h, hipGetLastError());
    double tot[4];
    HCHK(h, hipMemcpyAsync(tot, fin, sizeof(tot), hipMemcpyDeviceToHost, h->stream));
    HCHK(h, hipStreamSynchronize(h->stream));
    const double a = h->P.dx * h->P.dy;
    out->t = t_new;
    out->dt = s.dt;
    out->mass = a * tot[0];
    out->volume = a * tot[1];
    out->area = a * tot[2];
    out->solid_mass = a * tot[3];
    out->active_cells = (int64_t)s.active;
    out->dried_cells = (int64_t)s.dried;
    out->err = err;
  } else {
    HCHK(h, hipStreamSynchronize(h->stream));
  }
  return (int)(err & ~IMEX_WARN_RESCALE);
}

// Start of a device-driven call: the call's t_end, the end condition of the
// driver loop (t < t_end, src/IMEX_SfloW2D.f90:395) for its first step, and
// cleared physics error bits.
// The passes skipped after the previous call reached its t_end built no
// lists, so the next pass visits every tile.
static int begin_driven(imex_handle* h, double t_end, DevState* s) {
  if (read_state(h, s)) return -1;
  if (s->stop_next) h->mask_full_next = true;
  s->err = 0u;  // the call returns the bits its own steps raise
  s->t_end = t_end;
  s->done = s->stop_next = (s->t < t_end) ? 0u : 1u;
  HCHK(h, hipMemcpyAsync(h->G.st, s, sizeof(*s), hipMemcpyHostToDevice, h->stream));
  return 0;
}

extern "C" int imex_run_steps(imex_handle* h, int nsteps, double t_end, double* dts) {
  if (!h || nsteps < 0) return -2;
  if (int rc = need_topography(h)) return rc;
  HCHK(h, hipSetDevice(h->device));
  if (int rc = bind_dts(h, dts, nsteps)) return rc;
  {
    DevState s;
    if (int rc = begin_driven(h, t_end, &s)) return rc;
  }
  int rc;
  for (int n = 0; n < nsteps; n++) {
    if (h->P.general) {
      // the general path: every tile's active set, then imex_pore.hip's launches
      h->mask_full_next = true;
      if ((rc = mask(h, 0, 1))) return rc;
      if ((rc = imex_pore_step(h, n))) return rc;
      continue;
    }
    if ((rc = mask(h, 0, 1))) return rc;
    if ((rc = flux(h, 1, 0.0, n))) return rc;
    if ((rc = update(h, 1 | 2 | 4, 0.0, 1, n))) return rc;
    if ((rc = topo(h))) return rc;
  }
  if (nsteps > 0 && (rc = commit(h))) return rc;
  if (dts && nsteps > 0)
    HCHK(h, hipMemcpyAsync(dts, h->G.dts, 8 * (size_t)nsteps, hipMemcpyDeviceToHost, h->stream));
  unsigned err = 0;
  if (take_err(h, &err)) return -1;
  return (int)(err & ~IMEX_WARN_RESCALE);
}

extern "C" int imex_run_steps_timed(imex_handle* h, int nsteps, double t_end, double* dts,
                                    int sample_every, double* kernel_ms, int* n_sampled,
                                    int64_t* cells) {
  if (!h || nsteps < 0 || sample_every < 1) return -2;
  if (int rc = device_driven_only(h)) return rc;  // its kernel split is the step kernels'
  if (int rc = need_topography(h)) return rc;
  HCHK(h, hipSetDevice(h->device));
  if (int rc = bind_dts(h, dts, nsteps)) return rc;
  DevState s;
  if (int r = begin_driven(h, t_end, &s)) return r;
  const unsigned long long act0 = s.active;
  constexpr int NK = 3;  // k_mask, k_flux, k_update
  const int ns = (nsteps + sample_every - 1) / sample_every;  // steps n with n % sample_every == 0
  std::vector<hipEvent_t> ev((size_t)ns * (NK + 1), nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = set_err(h, "hipEventCreate", hipErrorOutOfMemory);
  // an event record costs a few microseconds of stream time: only the sampled
  // steps carry them, the others run as imex_run_steps does
  auto rec = [&](hipEvent_t* e, int k) {
    if (e && !rc) {
      const hipError_t x = hipEventRecord(e[k], h->stream);
      if (x != hipSuccess) rc = set_err(h, "hipEventRecord", x);
    }
  };
  for (int n = 0; n < nsteps && !rc; n++) {
    hipEvent_t* e = (n % sample_every == 0) ? &ev[(size_t)(NK + 1) * (n / sample_every)] : nullptr;
    rec(e, 0);
    if (!rc) rc = mask(h, 0, 1);
    rec(e, 1);
    if (!rc) rc = flux(h, 1, 0.0, n);
    rec(e, 2);
    if (!rc) rc = update(h, 1 | 2 | 4, 0.0, 1, n);
    if (!rc) rc = topo(h);
    rec(e, 3);
  }
  if (!rc && nsteps > 0) rc = commit(h);
  const hipError_t sy = hipStreamSynchronize(h->stream);
  if (!rc && sy != hipSuccess) rc = set_err(h, "hipStreamSynchronize", sy);
  double ms[NK] = {0, 0, 0};
  for (int m = 0; m < ns && !rc; m++)
    for (int k = 0; k < NK && !rc; k++) {
      float f = 0.f;
      const hipError_t x = hipEventElapsedTime(&f, ev[(size_t)(NK + 1) * m + k], ev[(size_t)(NK + 1) * m + k + 1]);
      if (x != hipSuccess) rc = set_err(h, "hipEventElapsedTime", x);
      ms[k] += f;
    }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc) return rc;
  if (kernel_ms)
    for (int k = 0; k < NK; k++) kernel_ms[k] = ms[k];
  if (n_sampled) *n_sampled = ns;
  if (dts && nsteps > 0)
    HCHK(h, hipMemcpyAsync(dts, h->G.dts, 8 * (size_t)nsteps, hipMemcpyDeviceToHost, h->stream));
  if (read_state(h, &s)) return -1;
  if (cells) cells[0] = (int64_t)(s.active - act0);
  return (int)(s.err & ~IMEX_WARN_RESCALE);
}

// ===========================================================================
// Batches: independent runs (ensemble members, served requests) advanced
// together.  One launch per kernel and step covers every member: blockIdx.y
// selects the member, whose constants and buffers come from a table in
// device memory, and the member's blocks walk its own lists exactly as the
// single-run kernels do, so a member's arithmetic -- and result -- is the
// same alone and in a batch.  Each member keeps its own dt, t and t_end; a
// member whose t has reached t_end does nothing.  A batch step is k_mask,
// k_flux, k_update<7> and a commit launch that closes the step of every
// member (dt history, t += dt, the done flag) before the next step's
// launches read them.
// ===========================================================================
struct Member {
  Params P;
  Grid G;
};

// The pass parity and number of the batch's next step, in device memory so
// that one captured step replays: the step's launches read them and its
// commit launch advances them.
struct BatchCtl {
  int par;
  unsigned int seq;
};
__global__ void kb_set_ctl(BatchCtl* c, int par, unsigned int seq) {
  c->par = par;
  c->seq = seq;
}

template <typename S>
__global__ __launch_bounds__(NTHR) void kb_mask(const Member* M, const BatchCtl* ctl) {
  const Member& m = M[blockIdx.y];
  const DevState* st = m.G.st;
  if (st->done) return;
  mask_body<S>(m.P, m.G, 0, ctl->par, ctl->seq, st->full_next ? 1 : 0);
}
template <int NS, typename F, int THR>
__global__ __launch_bounds__(THR) void kb_flux(const Member* M, const BatchCtl* ctl) {
  const Member& m = M[blockIdx.y];
  if (m.G.st->done) return;
  flux_body<NS, F, THR>(m.P, m.G, 1, 0.0, ctl->par);
}
template <int NS, typename S>
__global__ __launch_bounds__(NTHR) void kb_update(const Member* M) {
  const Member& m = M[blockIdx.y];
  if (m.G.st->done) return;
  update_body<NS, 7, S>(m.P, m.G, 0.0, 1, 0);
}
__global__ void kb_commit(const Member* M, int n, BatchCtl* ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    DevState* st = M[i].G.st;
    if (!st->done) {
      commit_step(st);
      st->full_next = 0u;
    }
  }
  if (i == 0) {
    ctl->par ^= 1;
    ctl->seq += 1u;
  }
}
__global__ void kb_status(const Member* M, int n, unsigned int* done, double* t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    done[2 * i] = M[i].G.st->done;
    done[2 * i + 1] = M[i].G.st->steps;
    t[i] = M[i].G.st->t;
  }
}

struct imex_batch {
  int ns = 0;
  int device = 0;
  int n_cu = 256;
  bool fp32 = false;                 // members of the FP32 form
  int gx = 0;                        // blocks per member and launch; 0: from the member count
  int flux_thr = NTHR;               // kb_flux block: NTHR (two passes per phase) or 2 NTHR
  bool graphs = true;                // replay captured steps (else launch kernel by kernel)
  BatchCtl* d_ctl = nullptr;
  // captured steps: one step and GSTEPS steps, for the launch shape in key
  hipGraphExec_t g1 = nullptr, gk = nullptr;
  hipGraph_t graph1 = nullptr, graphk = nullptr;
  long long key[4] = {0, 0, 0, 0};
  hipStream_t stream = nullptr;
  std::vector<imex_handle*> mem;
  Member* d_mem = nullptr;
  int cap = 0;
  bool dirty = true;
  int par = 0;
  unsigned int seq = 0;
  size_t lds = 0;
  unsigned int* d_done = nullptr;
  double* d_t = nullptr;
  int status_cap = 0;
  std::string err;
};

static int bset_err(imex_batch* b, const std::string& what, hipError_t e) {
  b->err = what + ": " + hipGetErrorString(e);
  return -1;
}
#define BCHK(b, call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return bset_err((b), #call, e_); \
  } while (0)

// Empties a run's tile, candidate and work lists and their stamps, so that
// they can follow another sequence of pass numbers (joining or leaving a
// batch); the next active-set pass visits every tile.  Sets t_end and the
// batch flags (done, full pass next).
static int reset_lists(imex_handle* h, double t_end, unsigned int full_next) {
  HCHK(h, hipMemsetAsync(h->G.cand_stamp, 0, sizeof(unsigned int) * (size_t)h->ntiles, h->stream));
  DevState s;
  if (read_state(h, &s)) return -1;
  s.n_tiles[0] = s.n_tiles[1] = 0u;
  s.n_cand[0] = s.n_cand[1] = s.n_cand[2] = 0u;
  s.n_work = 0u;
  s.t_end = t_end;
  s.done = (s.t < t_end) ? 0u : 1u;
  s.full_next = full_next;
  HCHK(h, hipMemcpyAsync(h->G.st, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  h->mask_full_next = true;
  h->mask_seq = 0;
  return 0;
}

// the kb_flux kernel of a batch's solid classes, precision and block size
static const void* kb_flux_fn(int ns, bool fp32, int thr) {
  if (thr == NTHR) {
    switch (ns) {
      case 1: return fp32 ? (const void*)kb_flux<1, float, NTHR> : (const void*)kb_flux<1, double, NTHR>;
      case 2: return fp32 ? (const void*)kb_flux<2, float, NTHR> : (const void*)kb_flux<2, double, NTHR>;
      default: return fp32 ? (const void*)kb_flux<3, float, NTHR> : (const void*)kb_flux<3, double, NTHR>;
    }
  }
  switch (ns) {
    case 1: return fp32 ? (const void*)kb_flux<1, float, FLUX_THR> : (const void*)kb_flux<1, double, FLUX_THR>;
    case 2: return fp32 ? (const void*)kb_flux<2, float, FLUX_THR> : (const void*)kb_flux<2, double, FLUX_THR>;
    default: return fp32 ? (const void*)kb_flux<3, float, FLUX_THR> : (const void*)kb_flux<3, double, FLUX_THR>;
  }
}
static const void* kb_update_fn(int ns, bool fp32) {
  switch (ns) {
    case 1: return fp32 ? (const void*)kb_update<1, float> : (const void*)kb_update<1, double>;
    case 2: return fp32 ? (const void*)kb_update<2, float> : (const void*)kb_update<2, double>;
    default: return fp32 ? (const void*)kb_update<3, float> : (const void*)kb_update<3, double>;
  }
}

static void batch_drop_graphs(imex_batch* b) {
  if (b->g1) (void)hipGraphExecDestroy(b->g1);
  if (b->gk) (void)hipGraphExecDestroy(b->gk);
  if (b->graph1) (void)hipGraphDestroy(b->graph1);
  if (b->graphk) (void)hipGraphDestroy(b->graphk);
  b->g1 = b->gk = nullptr;
  b->graph1 = b->graphk = nullptr;
}

extern "C" int imex_batch_create_form(int n_solid, int precision, imex_batch** out) {
  if (!out || n_solid < 1 || n_solid > 3 || precision < 0 || precision > 1) return -2;
  imex_batch* b = new imex_batch();
  *out = b;
  b->ns = n_solid;
  b->fp32 = precision == 1;
  const char* dev = getenv("IMEX_DEVICE");
  b->device = dev ? atoi(dev) : 0;
  BCHK(b, hipSetDevice(b->device));
  BCHK(b, hipStreamCreate(&b->stream));
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && ncu > 0)
    b->n_cu = ncu;
  for (int thr : {NTHR, FLUX_THR})
    BCHK(b, hipFuncSetAttribute(kb_flux_fn(n_solid, b->fp32, thr), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_for(n_solid, b->fp32, thr == NTHR)));
  b->lds = lds_for(n_solid, b->fp32, true);
  BCHK(b, hipMalloc(&b->d_ctl, sizeof(BatchCtl)));
  // the first step's pass parity and number
  b->par = 0;
  b->seq = 0;
  hipLaunchKernelGGL(kb_set_ctl, dim3(1), dim3(1), 0, b->stream, b->d_ctl, 1, 1u);
  BCHK(b, hipGetLastError());
  BCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

extern "C" int imex_batch_create(int n_solid, imex_batch** out) {
  return imex_batch_create_form(n_solid, 0, out);
}

extern "C" int imex_batch_set_blocks(imex_batch* b, int blocks_per_member) {
  if (!b || blocks_per_member < 0 || blocks_per_member > 4096) return -2;
  b->gx = blocks_per_member;
  return 0;
}

extern "C" int imex_batch_set_launch(imex_batch* b, int flux_threads, int graphs) {
  if (!b || (flux_threads != NTHR && flux_threads != FLUX_THR) || graphs < 0 || graphs > 1) return -2;
  b->flux_thr = flux_threads;
  b->lds = lds_for(b->ns, b->fp32, flux_threads == NTHR);
  b->graphs = graphs == 1;
  return 0;
}

extern "C" int imex_batch_add(imex_batch* b, imex_handle* h, double t_end) {
  // members step together with no per-member slope pass: no topography change
  if (!b || !h || h->P.ns != b->ns || h->device != b->device || h->fp32 != b->fp32 ||
      h->P.topo_change || h->P.general)
    return -2;
  for (imex_handle* x : b->mem)
    if (x == h) return -2;
  // the member's own stream is idle from here on; its state joins the batch
  // with a full active-set pass at the batch's next step
  if (commit(h)) {
    b->err = h->err;
    return -1;
  }
  if (reset_lists(h, t_end, 1u)) {
    b->err = h->err;
    return -1;
  }
  b->mem.push_back(h);
  b->dirty = true;
  return 0;
}

extern "C" int imex_batch_remove(imex_batch* b, imex_handle* h) {
  if (!b || !h) return -2;
  for (size_t i = 0; i < b->mem.size(); i++)
    if (b->mem[i] == h) {
      BCHK(b, hipStreamSynchronize(b->stream));
      b->mem.erase(b->mem.begin() + (long)i);
      b->dirty = true;
      // back to its own pass counters: fresh lists, a full pass first
      DevState s;
      if (read_state(h, &s) || reset_lists(h, s.t_end, 0u)) {
        b->err = h->err;
        return -1;
      }
      return 0;
    }
  return -2;
}

static int batch_table(imex_batch* b) {
  if (!b->dirty) return 0;
  const int n = (int)b->mem.size();
  if (n > b->cap) {
    BCHK(b, hipStreamSynchronize(b->stream));
    if (b->d_mem) BCHK(b, hipFree(b->d_mem));
    b->d_mem = nullptr;
    const int cap = std::max(n, 2 * b->cap);
    BCHK(b, hipMalloc(&b->d_mem, sizeof(Member) * (size_t)cap));
    b->cap = cap;
  }
  std::vector<Member> host((size_t)n);
  for (int i = 0; i < n; i++) {
    host[i].P = b->mem[i]->P;
    host[i].G = b->mem[i]->G;
    host[i].G.dts = nullptr;
  }
  if (n > 0) {
    BCHK(b, hipMemcpyAsync(b->d_mem, host.data(), sizeof(Member) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));  // the host vector goes out of scope
  }
  b->dirty = false;
  return 0;
}

// One batch step: the active sets, the fluxes and the updates of every
// member, then the commit (which also advances the step's pass parity and
// number in BatchCtl).  Kernel arguments stay the same from step to step.
static int batch_launch_step(imex_batch* b, int n, int gx) {
  const dim3 g(gx, n), blk(NTHR);
  if (b->fp32)
    hipLaunchKernelGGL(kb_mask<float>, g, blk, 0, b->stream, b->d_mem, b->d_ctl);
  else
    hipLaunchKernelGGL(kb_mask<double>, g, blk, 0, b->stream, b->d_mem, b->d_ctl);
  const Member* m = b->d_mem;
  const BatchCtl* c = b->d_ctl;
  void* fargs[] = {(void*)&m, (void*)&c};
  BCHK(b, hipLaunchKernel(kb_flux_fn(b->ns, b->fp32, b->flux_thr), g, dim3(b->flux_thr), fargs, b->lds,
                          b->stream));
  void* uargs[] = {(void*)&m};
  BCHK(b, hipLaunchKernel(kb_update_fn(b->ns, b->fp32), g, blk, uargs, 0, b->stream));
  hipLaunchKernelGGL(kb_commit, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_mem, n, b->d_ctl);
  BCHK(b, hipGetLastError());
  return 0;
}

constexpr int GSTEPS = 8;  // steps in the longer captured graph

// captures k steps of the current launch shape into an executable graph
static int batch_capture(imex_batch* b, int n, int gx, int k, hipGraph_t* graph, hipGraphExec_t* exec) {
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int i = 0; i < k && !rc; i++) rc = batch_launch_step(b, n, gx);
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return bset_err(b, "hipStreamEndCapture", e);
  *graph = gr;
  BCHK(b, hipGraphInstantiate(exec, gr, nullptr, nullptr, 0));
  return 0;
}

// nsteps batch steps; members whose t reaches t_end stop on the way.
extern "C" int imex_batch_run(imex_batch* b, int nsteps) {
  if (!b || nsteps < 0) return -2;
  BCHK(b, hipSetDevice(b->device));
  if (int rc = batch_table(b)) return rc;
  const int n = (int)b->mem.size();
  if (n == 0 || nsteps == 0) return 0;
  if (n > 65535) return -2;
  // automatic: about 16 blocks per CU in all, at least 8 and at most 1024 per member
  const int gx = b->gx ? b->gx : std::min(1024, std::max(8, 16 * b->n_cu / n));
  if (!b->graphs) {
    for (int k = 0; k < nsteps; k++)
      if (int rc = batch_launch_step(b, n, gx)) return rc;
  } else {
    const long long key[4] = {(long long)(uintptr_t)b->d_mem, n, gx, b->flux_thr};
    if (!std::equal(key, key + 4, b->key)) {
      BCHK(b, hipStreamSynchronize(b->stream));
      batch_drop_graphs(b);
      std::copy(key, key + 4, b->key);
    }
    if (!b->g1)
      if (int rc = batch_capture(b, n, gx, 1, &b->graph1, &b->g1)) return rc;
    if (nsteps >= GSTEPS && !b->gk)
      if (int rc = batch_capture(b, n, gx, GSTEPS, &b->graphk, &b->gk)) return rc;
    int k = 0;
    for (; k + GSTEPS <= nsteps; k += GSTEPS) BCHK(b, hipGraphLaunch(b->gk, b->stream));
    for (; k < nsteps; k++) BCHK(b, hipGraphLaunch(b->g1, b->stream));
  }
  // the host's copy of the pass parity and number (the device's advance per step)
  b->par ^= nsteps & 1;
  b->seq += (unsigned int)nsteps;
  return 0;
}

// done flag and t of every member, in the order they were added (members
// removed since are not listed); synchronises the batch stream
extern "C" int imex_batch_status(imex_batch* b, int n, int32_t* done, int32_t* steps, double* t) {
  if (!b || n != (int)b->mem.size()) return -2;
  if (n == 0) return 0;
  BCHK(b, hipSetDevice(b->device));
  if (int rc = batch_table(b)) return rc;
  if (n > b->status_cap) {
    BCHK(b, hipStreamSynchronize(b->stream));
    if (b->d_done) BCHK(b, hipFree(b->d_done));
    if (b->d_t) BCHK(b, hipFree(b->d_t));
    b->d_done = nullptr;
    b->d_t = nullptr;
    BCHK(b, hipMalloc(&b->d_done, 2 * sizeof(unsigned int) * (size_t)n));
    BCHK(b, hipMalloc(&b->d_t, sizeof(double) * (size_t)n));
    b->status_cap = n;
  }
  hipLaunchKernelGGL(kb_status, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_mem, n, b->d_done, b->d_t);
  BCHK(b, hipGetLastError());
  std::vector<unsigned int> dn(2 * (size_t)n);
  BCHK(b, hipMemcpyAsync(dn.data(), b->d_done, 2 * sizeof(unsigned int) * (size_t)n, hipMemcpyDeviceToHost, b->stream));
  if (t) BCHK(b, hipMemcpyAsync(t, b->d_t, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; i++) {
    if (done) done[i] = (int32_t)dn[2 * i];
    if (steps) steps[i] = (int32_t)dn[2 * i + 1];
  }
  return 0;
}

extern "C" int imex_batch_last_error(imex_batch* b, char* buf, size_t n) {
  if (!b || !buf || n == 0) return -2;
  const size_t k = std::min(n - 1, b->err.size());
  memcpy(buf, b->err.data(), k);
  buf[k] = 0;
  return 0;
}

extern "C" void imex_batch_destroy(imex_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graphs(b);
  if (b->d_mem) (void)hipFree(b->d_mem);
  if (b->d_ctl) (void)hipFree(b->d_ctl);
  if (b->d_done) (void)hipFree(b->d_done);
  if (b->d_t) (void)hipFree(b->d_t);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

extern "C" int imex_calib_copy(imex_handle* h, int64_t n, int reps) {
  if (!h || n <= 0 || reps <= 0) return -2;
  HCHK(h, hipSetDevice(h->device));
  double *a = nullptr, *b = nullptr;
  HCHK(h, hipMalloc(&a, 8 * (size_t)n));
  hipError_t e = hipMalloc(&b, 8 * (size_t)n);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return set_err(h, "hipMalloc", e);
  }
  HCHK(h, hipMemsetAsync(a, 0, 8 * (size_t)n, h->stream));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, h->stream, a, b, (long long)n);
  e = hipStreamSynchronize(h->stream);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return set_err(h, "k_copy", e);
  return 0;
}

extern "C" int imex_get_counters(imex_handle* h, int64_t* active_cells, int64_t* steps) {
  if (!h) return -2;
  DevState s;
  if (read_state(h, &s)) return -1;
  if (active_cells) *active_cells = (int64_t)s.active;
  if (steps) *steps = (int64_t)s.steps;
  return 0;
}

extern "C" int imex_get_footprint(imex_handle* h, size_t* bytes) {
  if (!h || !bytes) return -2;
  size_t b = 0;
  for (size_t x : h->alloc_sizes) b += x;
  *bytes = b;
  return 0;
}

extern "C" int imex_get_work_counters(imex_handle* h, int64_t* work_cells, int64_t* frame_cells) {
  if (!h) return -2;
  DevState s;
  if (read_state(h, &s)) return -1;
  if (work_cells) *work_cells = (int64_t)s.work;
  if (frame_cells) *frame_cells = (int64_t)h->G.static_active;
  return 0;
}

extern "C" int imex_get_time_history(imex_handle* h, double* t, double* dt_old,
                                     double* dt_old_old) {
  if (!h) return -2;
  DevState s;
  if (read_state(h, &s)) return -1;
  if (t) *t = s.t;
  if (dt_old) *dt_old = s.dt_old;
  if (dt_old_old) *dt_old_old = s.dt_old_old;
  return 0;
}

extern "C" int imex_get_time(imex_handle* h, double* t, double* dt) {
  if (!h) return -2;
  DevState s;
  if (read_state(h, &s)) return -1;
  if (t) *t = s.t;
  if (dt) *dt = s.dt;
  return 0;
}

extern "C" int imex_set_time(imex_handle* h, double t, double dt_old, double dt_old_old) {
  if (!h) return -2;
  HCHK(h, hipSetDevice(h->device));
  DevState s;
  if (read_state(h, &s)) return -1;
  s.t = t;
  s.dt_old = dt_old;
  s.dt_old_old = dt_old_old;
  HCHK(h, hipMemcpyAsync(h->G.st, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int imex_download(imex_handle* h, unsigned what, double* q_aos, double* hmax,
                             double* pdynmax, double* modvelmax, double* deposit, int32_t* vuln,
                             int32_t* maskout) {
  if (!h) return -2;
  HCHK(h, hipSetDevice(h->device));
  const int N = h->G.N, nv = h->P.nv, ns = h->P.ns;
  if ((what & IMEX_DL_Q) && q_aos) {
    if (h->fp32)
      hipLaunchKernelGGL(k_soa_to_aos<float>, dim3((N + 255) / 256), dim3(256), 0, h->stream, h->G.qf, h->d_tmp, nv, N);
    else
      hipLaunchKernelGGL(k_soa_to_aos<double>, dim3((N + 255) / 256), dim3(256), 0, h->stream, h->G.q, h->d_tmp, nv, N);
    HCHK(h, hipGetLastError());
    HCHK(h, hipMemcpyAsync(q_aos, h->d_tmp, 8 * (size_t)N * nv, hipMemcpyDeviceToHost, h->stream));
  }
  if (what & IMEX_DL_STATS) {
    if (hmax) HCHK(h, hipMemcpyAsync(hmax, h->G.hmax, 8 * (size_t)N, hipMemcpyDeviceToHost, h->stream));
    if (pdynmax) HCHK(h, hipMemcpyAsync(pdynmax, h->G.pdyn, 8 * (size_t)N, hipMemcpyDeviceToHost, h->stream));
    if (modvelmax) HCHK(h, hipMemcpyAsync(modvelmax, h->G.mvel, 8 * (size_t)N, hipMemcpyDeviceToHost, h->stream));
  }
  if ((what & IMEX_DL_DEPOSIT) && deposit)
    HCHK(h, hipMemcpyAsync(deposit, h->G.dep, 8 * (size_t)N * ns, hipMemcpyDeviceToHost, h->stream));
  std::vector<unsigned long long> vb;
  std::vector<unsigned char> mb;
  if ((what & IMEX_DL_VULN) && vuln && h->G.vuln) {
    vb.resize(N);
    HCHK(h, hipMemcpyAsync(vb.data(), h->G.vuln, 8 * (size_t)N, hipMemcpyDeviceToHost, h->stream));
  }
  if ((what & IMEX_DL_MASK) && maskout) {
    mb.resize(N);
    HCHK(h, hipMemcpyAsync(mb.data(), h->G.mask, (size_t)N, hipMemcpyDeviceToHost, h->stream));
  }
  HCHK(h, hipStreamSynchronize(h->stream));
  if (!vb.empty()) {
    const int nl = h->P.n_thk * h->P.n_pdyn;
    for (int c = 0; c < N; c++)
      for (int l = 0; l < nl; l++) vuln[(size_t)c * nl + l] = (int32_t)((vb[c] >> l) & 1ULL);
  }
  if (!mb.empty())
    for (int c = 0; c < N; c++) maskout[c] = mb[c] & 1;
  return 0;
}

extern "C" int imex_topography_reconstruction(imex_handle* h) {
  if (!h) return -2;
  if (int rc = need_topography(h)) return rc;
  HCHK(h, hipSetDevice(h->device));
  return topo(h);
}

extern "C" int imex_set_erodible(imex_handle* h, const double* erodible) {
  if (!h || !erodible) return -2;
  if (!h->G.erodible) {
    h->err = "imex_set_erodible: the run has no erosion (erosion_coeff = 0)";
    return -2;
  }
  HCHK(h, hipSetDevice(h->device));
  const size_t N = (size_t)h->G.N;
  const int ns = h->P.ns;
  // erodible(n_solid, nx, ny) -> [class][cell]
  std::vector<double> t(N * ns);
  for (size_t c = 0; c < N; c++)
    for (int i = 0; i < ns; i++) t[(size_t)i * N + c] = erodible[c * ns + i];
  HCHK(h, hipMemcpyAsync(h->G.erodible, t.data(), 8 * N * ns, hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int imex_set_noise(imex_handle* h, const double* Z) {
  if (!h || !Z) return -2;
  if (!h->G.noise) {
    h->err = "imex_set_noise: the run has no stochastic friction (stochastic_flag = 0)";
    return -2;
  }
  HCHK(h, hipSetDevice(h->device));
  HCHK(h, hipMemcpyAsync(h->G.noise, Z, 8 * (size_t)h->G.N, hipMemcpyHostToDevice, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int imex_get_erosion(imex_handle* h, double* erosion, double* erodible) {
  if (!h) return -2;
  HCHK(h, hipSetDevice(h->device));
  const size_t N = (size_t)h->G.N;
  const int ns = h->P.ns;
  if (!h->G.erodible) {
    if (erosion) std::fill(erosion, erosion + N * ns, 0.0);
    if (erodible) std::fill(erodible, erodible + N * ns, 0.0);
    return 0;
  }
  if (erosion)
    HCHK(h, hipMemcpyAsync(erosion, h->G.erosion, 8 * N * ns, hipMemcpyDeviceToHost, h->stream));
  std::vector<double> t(erodible ? N * ns : 0);
  if (erodible)
    HCHK(h, hipMemcpyAsync(t.data(), h->G.erodible, 8 * N * ns, hipMemcpyDeviceToHost, h->stream));
  HCHK(h, hipStreamSynchronize(h->stream));
  if (erodible)
    for (size_t c = 0; c < N; c++)
      for (int i = 0; i < ns; i++) erodible[c * ns + i] = t[(size_t)i * N + c];
  return 0;
}

extern "C" int imex_last_error(imex_handle* h, char* buf, size_t len) {
  if (!h || !buf || len == 0) return -2;
  snprintf(buf, len, "%s", h->err.c_str());
  return 0;
}

extern "C" void imex_destroy(imex_handle* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  imex_output_release(h);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

// ---- per-function batch entry points -------------------------------------
// Device buffers of one test-hook call; `bad` is set when an allocation or
// copy failed, and the hook then returns before launching anything.
struct DevBufs {
  imex_handle* h;
  std::vector<void*> p;
  bool bad = false;
  ~DevBufs() {
    for (void* x : p) (void)hipFree(x);
  }
  double* up(const double* src, size_t n) {
    void* d = nullptr;
    if (hipMalloc(&d, n ? 8 * n : 16) != hipSuccess) {
      bad = true;
      return nullptr;
    }
    p.push_back(d);
    if (src && hipMemcpy(d, src, 8 * n, hipMemcpyHostToDevice) != hipSuccess) bad = true;
    return (double*)d;
  }
  int* upi(size_t n) {
    void* d = nullptr;
    if (hipMalloc(&d, n ? 4 * n : 16) != hipSuccess) {
      bad = true;
      return nullptr;
    }
    p.push_back(d);
    return (int*)d;
  }
};
#define BUFS_OK(h, b) \
  do {                \
    if ((b).bad) {     \
      (h)->err = "test hook: device buffer allocation or upload failed"; \
      return -1;      \
    }                 \
  } while (0)

#define NSDISPATCH(h, KER, ...)                                                             \
  do {                                                                                      \
    const int nb_ = (n + 127) / 128;                                                        \
    switch ((h)->P.ns) {                                                                    \
      case 1: hipLaunchKernelGGL(KER<1>, dim3(nb_), dim3(128), 0, 0, (h)->P, __VA_ARGS__); break; \
      case 2: hipLaunchKernelGGL(KER<2>, dim3(nb_), dim3(128), 0, 0, (h)->P, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL(KER<3>, dim3(nb_), dim3(128), 0, 0, (h)->P, __VA_ARGS__); break; \
    }                                                                                       \
  } while (0)

static int fin(imex_handle* h) {
  HCHK(h, hipGetLastError());
  HCHK(h, hipDeviceSynchronize());
  return 0;
}
static int down(imex_handle* h, double* dst, const double* src, size_t n) {
  HCHK(h, hipMemcpy(dst, src, 8 * n, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int imex_test_qc_to_qp(imex_handle* h, int n, const double* qc, double* qp, double* pdyn) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *dq = b.up(qc, (size_t)n * nv), *dqp = b.up(nullptr, (size_t)n * (nv + 2)), *dp = b.up(nullptr, n);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_qc_to_qp, n, dq, dqp, dp);
  if (fin(h)) return -1;
  if (down(h, qp, dqp, (size_t)n * (nv + 2)) || down(h, pdyn, dp, n)) return -1;
  return 0;
}
extern "C" int imex_test_qp_to_qc(imex_handle* h, int n, const double* qp, double* qc) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *dqp = b.up(qp, (size_t)n * (nv + 2)), *dq = b.up(nullptr, (size_t)n * nv);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_qp_to_qc, n, dqp, dq);
  if (fin(h)) return -1;
  return down(h, qc, dq, (size_t)n * nv);
}
extern "C" int imex_test_fluxes(imex_handle* h, int n, int dir, const double* qc, const double* qp,
                                double* flux) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *dq = b.up(qc, (size_t)n * nv), *dqp = b.up(qp, (size_t)n * (nv + 2)), *df = b.up(nullptr, (size_t)n * nv);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_fluxes, n, dir, dq, dqp, df);
  if (fin(h)) return -1;
  return down(h, flux, df, (size_t)n * nv);
}
extern "C" int imex_test_speeds(imex_handle* h, int n, int dir, const double* qp, double* out2) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *dqp = b.up(qp, (size_t)n * (nv + 2)), *d = b.up(nullptr, (size_t)n * 2);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_speeds, n, dir, dqp, d);
  if (fin(h)) return -1;
  return down(h, out2, d, (size_t)n * 2);
}
extern "C" int imex_test_implicit_cplx(imex_handle* h, int n, const double* qre, const double* qim,
                                       double* ore, double* oim) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *a = b.up(qre, (size_t)n * nv), *c = b.up(qim, (size_t)n * nv);
  double *o1 = b.up(nullptr, (size_t)n * nv), *o2 = b.up(nullptr, (size_t)n * nv);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_implicit_cplx, n, a, c, o1, o2);
  if (fin(h)) return -1;
  if (down(h, ore, o1, (size_t)n * nv) || down(h, oim, o2, (size_t)n * nv)) return -1;
  return 0;
}
extern "C" int imex_test_implicit_real(imex_handle* h, int n, const double* q, double* out,
                                       int* fast) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *a = b.up(q, (size_t)n * nv), *o = b.up(nullptr, (size_t)n * nv);
  int* f = b.upi((size_t)n);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_implicit_real, n, a, o, f);
  if (fin(h)) return -1;
  if (down(h, out, o, (size_t)n * nv)) return -1;
  HCHK(h, hipMemcpy(fast, f, 4 * (size_t)n, hipMemcpyDeviceToHost));
  return 0;
}
extern "C" int imex_test_semi_impl(imex_handle* h, int n, const double* qp, double* si) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *dqp = b.up(qp, (size_t)n * (nv + 2)), *d = b.up(nullptr, (size_t)n * nv);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_semi_impl, n, dqp, d);
  if (fin(h)) return -1;
  return down(h, si, d, (size_t)n * nv);
}
extern "C" int imex_test_expl(imex_handle* h, int n, const double* bx, const double* by, double t,
                              const double* qp, double* expl) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *dbx = b.up(bx, n), *dby = b.up(by, n), *dqp = b.up(qp, (size_t)n * (nv + 2));
  double* d = b.up(nullptr, (size_t)n * nv);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_expl, n, dbx, dby, t, dqp, d);
  if (fin(h)) return -1;
  return down(h, expl, d, (size_t)n * nv);
}
extern "C" int imex_test_settling(imex_handle* h, int n, const double* diam, const double* rhos,
                                  const double* rhoc, const double* ikv, double* out) {
  if (!h || n <= 0) return -2;
  DevBufs b{h};
  double *a = b.up(diam, n), *c = b.up(rhos, n), *d = b.up(rhoc, n), *e = b.up(ikv, n), *o = b.up(nullptr, n);
  BUFS_OK(h, b);
  hipLaunchKernelGGL(kt_settling, dim3((n + 127) / 128), dim3(128), 0, 0, h->P, n, a, c, d, e, o);
  if (fin(h)) return -1;
  return down(h, out, o, n);
}
extern "C" int imex_test_mass_exchange(imex_handle* h, int n, const double* qp, double dt,
                                       double* dep, double* eq) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv, ns = h->P.ns;
  DevBufs b{h};
  double *dqp = b.up(qp, (size_t)n * (nv + 2)), *dd = b.up(nullptr, (size_t)n * ns), *de = b.up(nullptr, (size_t)n * nv);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_mass_exchange, n, dqp, (const double*)nullptr, dt, dd, de, (double*)nullptr);
  if (fin(h)) return -1;
  if (down(h, dep, dd, (size_t)n * ns) || down(h, eq, de, (size_t)n * nv)) return -1;
  return 0;
}
extern "C" int imex_test_mass_exchange_ers(imex_handle* h, int n, const double* qp,
                                           const double* erod, double dt, double* dep, double* eq,
                                           double* ers) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv, ns = h->P.ns;
  DevBufs b{h};
  double *dqp = b.up(qp, (size_t)n * (nv + 2)), *der = b.up(erod, (size_t)n * ns),
         *dd = b.up(nullptr, (size_t)n * ns), *de = b.up(nullptr, (size_t)n * nv),
         *drs = b.up(nullptr, (size_t)n * ns);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_mass_exchange, n, dqp, (const double*)der, dt, dd, de, drs);
  if (fin(h)) return -1;
  if (down(h, dep, dd, (size_t)n * ns) || down(h, eq, de, (size_t)n * nv) ||
      down(h, ers, drs, (size_t)n * ns))
    return -1;
  return 0;
}
extern "C" int imex_test_newton(imex_handle* h, int n, double dt, const double* q_si,
                                const double* q0, const double* R, double* q, int32_t* iters) {
  if (!h || n <= 0) return -2;
  const int nv = h->P.nv;
  DevBufs b{h};
  double *a = b.up(q_si, (size_t)n * nv), *c = b.up(q0, (size_t)n * nv), *r = b.up(R, (size_t)n * nv);
  double* o = b.up(nullptr, (size_t)n * nv);
  int* it = b.upi(n);
  BUFS_OK(h, b);
  NSDISPATCH(h, kt_newton, n, dt, a, c, r, o, it);
  if (fin(h)) return -1;
  if (down(h, q, o, (size_t)n * nv)) return -1;
  HCHK(h, hipMemcpy(iters, it, 4 * (size_t)n, hipMemcpyDeviceToHost));
  return 0;
}
extern "C" int imex_test_pow(imex_handle* h, int n, const double* x, const double* y, double* out) {
  if (!h || n <= 0) return -2;
  DevBufs b{h};
  double *dx = b.up(x, n), *dy = b.up(y, n), *o = b.up(nullptr, n);
  BUFS_OK(h, b);
  hipLaunchKernelGGL(kt_pow, dim3((n + 127) / 128), dim3(128), 0, 0, n, dx, dy, o);
  if (fin(h)) return -1;
  return down(h, out, o, (size_t)n);
}
extern "C" int imex_host_pow(int n, const double* x, const double* y, double* out) {
  if (n < 0 || (n > 0 && (!x || !y || !out))) return -2;
  for (int i = 0; i < n; i++) out[i] = pow_cr(x[i], y[i]);
  return 0;
}
extern "C" int imex_test_arith(imex_handle* h, int n, const double* a, const double* bb, double* out) {
  if (!h || n <= 0) return -2;
  DevBufs b{h};
  double *da = b.up(a, n), *db = b.up(bb, n), *o = b.up(nullptr, (size_t)4 * n);
  BUFS_OK(h, b);
  hipLaunchKernelGGL(kt_arith, dim3((n + 127) / 128), dim3(128), 0, 0, n, da, db, o);
  if (fin(h)) return -1;
  return down(h, out, o, (size_t)4 * n);
}

#ifdef IMEX_PHASE_CLOCKS
// phase-clock build only (not part of include/imex_hip.h): switch the record
// on for the following update launches, and read it
extern "C" int imex_phase_clocks_enable(int on) {
  return hipMemcpyToSymbol(HIP_SYMBOL(g_phase_on), &on, sizeof(int)) == hipSuccess ? 0 : -1;
}
extern "C" int imex_phase_clocks_read(unsigned long long* out, int n_cells) {
  if (!out || n_cells < 0) return -2;
  const size_t n = (size_t)std::min(n_cells, kPhaseCells) * PC_NSLOT;
  return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase), n * sizeof(unsigned long long)) == hipSuccess
             ? 0 : -1;
}
#endif
