// Kernels of the FP64 IMEX-RK stage on gfx950.
//
// Per time step (one HIP stream, no host synchronisation in the device-side
// driver loop):
//   k_mask    active set (check_solve): seeds = wet cells, ring cells and the
//             boundary frame, dilated by the diamond of radius n_RK = 2; one
//             byte per cell.  Visits the candidate tiles the previous pass
//             listed (hot tiles and their neighbours) -- every other tile
//             holds only its frame band -- and lists the "hot" active half
//             tiles (a wet or source-ring cell within two cells) for k_flux.
//   k_flux    one 512-thread block per CU over the hot half tiles (one half
//             tile per block in one pass per phase, or two per block on its
//             two 256-thread halves when there are more half tiles than
//             blocks; batch members: 256-thread blocks); per
//             FX x TY half tile with a two-cell halo in LDS: primitive states -> MUSCL/minmod face
//             states -> local speeds -> Kurganov-Tadmor face fluxes -> flux
//             divergence minus explicit terms (the stage-1 right-hand side),
//             the CFL candidates reduced per block into one atomic minimum,
//             and the list of work cells (active cells whose state or
//             right-hand side is not all zero).
//   k_dt      one thread, host-driven steps only: the CFL time step.
//   k_update  persistent blocks over the work cells, an aligned lane pair
//             per cell: stage-2 semi-implicit Coulomb friction, Newton solve
//             of the turbulent friction (complex-step Jacobian columns and
//             line-search trials on the pair, 2x2 Cramer), assembly checks,
//             sedimentation / entrainment (settling classes on the pair) and
//             the driver epilogue (running maxima, vulnerability table).  The
//             device-side driver runs all of it in one launch <7> and takes
//             the step's dt (clip to t_end, growth limit) at its start; the
//             next launch (k_mask, or k_commit after the last step) commits
//             the dt history and t += dt.  Host-driven steps launch the parts
//             <1>, <2>, <4> from imex_rk_step, imex_erosion_deposition and
//             imex_post_step.
// Lists are double- or triple-buffered by pass parity so that no kernel waits
// on another through device-scope fences: each launch only clears counters of
// slots no block of that launch reads or fills.
#pragma once

#include "imex_physics.h"

namespace imex {

constexpr int TX = 32;  // tile width (cells, x fastest)
constexpr int TY = 8;   // tile height
constexpr int HALO = 2;
constexpr int RX = TX + 2 * HALO;
constexpr int RY = TY + 2 * HALO;
constexpr int NTHR = 256;
// k_flux works on half tiles (FX x TY cells): the halo states, the x faces,
// the y faces and the cells then each take one pass of the NTHR threads
constexpr int FX = TX / 2;
constexpr int FRX = FX + 2 * HALO;

// device-side driver state (one per handle)
struct DevState {
  double t, dt, dt_old, dt_old_old, t_end, dt_cfl;
  unsigned long long dtmin_bits;
  unsigned int n_tiles[2];  // hot half tiles listed by k_mask, per list slot
  unsigned int err;
  unsigned int n_work;      // cells k_flux listed for k_update
  unsigned int n_cand[3];   // candidate tiles of k_mask passes, three slots
  unsigned int pending;      // 1: the last step's dt history and t are still to commit
  unsigned int done;         // batch member: t reached t_end, its kernels do nothing
  unsigned int full_next;    // batch member: the next active-set pass visits every tile
  unsigned int steps;        // device-driven steps committed since the upload
  unsigned int stop_next;    // single run, device-driven: the step taken last reaches t_end,
                             // so the following active-set passes do no tile work
  unsigned int hot;          // set by a .q_2d snapshot that left a dry cell with other
                             // non-zero components: every active-set pass then visits
                             // every tile that holds one (cleared by an upload)
  unsigned long long dried;
  unsigned long long active;
  unsigned long long work;   // work-list cells summed over the stage-2 launches since create
};

struct Grid {
  double* q;     // [nv][N] conservative state (FP64 path)
  double* R;     // [nv][N] stage-1 right-hand side divF - expl (FP64 path)
  float* qf;     // [nv][N] the state in the FP32 form (q unused there)
  float* Rf;     // [nv][N] the right-hand side in the FP32 form
  double* hmax;  // [N]
  double* pdyn;  // [N]
  double* mvel;  // [N]
  double* dep;   // [ns][N]
  double* erodible;  // [ns][N] erodible layer (erosion runs only)
  double* erosion;   // [ns][N] eroded thickness (erosion runs only)
  unsigned long long* vuln;  // [N] bit (i_thk*n_pdyn + i_pdyn)
  unsigned char* mask;       // [N]
  int* tile_list;            // [2][2 ntiles] hot half tiles (2 tile + half), two slots
  int* work;                 // [N] active cells with a non-zero state or right-hand side
  int* cand;                 // [3][ntiles] tiles the next k_mask pass visits
  unsigned int* cand_stamp;  // [ntiles] pass that last listed the tile
  int* tile_static;          // [ntiles] frame-band active cells of the tile
  unsigned long long static_active;  // their sum
  double* bx;                // [N] B_prime_x_geom (recomputed each step with topography change)
  double* by;                // [N] B_prime_y_geom
  double* bcent;             // [N] B_cent, once the caller set it (outputs, topography change)
  double* noise;             // [N] stochastic value Z of each cell (stochastic friction only)
  const signed char* scell;  // [N] source_cell
  const unsigned char* sides;  // [N] sourceE/W/N/S bits
  const double* arcp;        // [N]
  const double* arcx;        // [N]
  const double* arcy;        // [N]
  const double* xc;          // [nx]   x_comp
  const double* xs;          // [nx+1] x_stag
  const double* yc;          // [ny]
  const double* ys;          // [ny+1]
  DevState* st;
  double* dts;               // per-step dt record of the device-side driver
  int nx, ny, N, ntx, nty;
};

// the state and right-hand side arrays of the storage type S
template <typename S>
__device__ __forceinline__ S* state_q(const Grid& G);
template <>
__device__ __forceinline__ double* state_q<double>(const Grid& G) { return G.q; }
template <>
__device__ __forceinline__ float* state_q<float>(const Grid& G) { return G.qf; }
template <typename S>
__device__ __forceinline__ S* state_R(const Grid& G);
template <>
__device__ __forceinline__ double* state_R<double>(const Grid& G) { return G.R; }
template <>
__device__ __forceinline__ float* state_R<float>(const Grid& G) { return G.Rf; }

// Block index -> tile index so that each XCD (blocks are dispatched round-robin
// over the 8 XCDs) works on a contiguous band of tiles and finds its halo
// neighbours in its own L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int x = b & 7, slot = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + slot;  // a bijection of 0..nb-1
}

}  // namespace imex
