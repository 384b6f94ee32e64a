// Host-side handle of one run and the helpers the C-ABI translation units
// share (imex_hip.hip: the step; imex_output.hip: the outputs).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "imex_hip.h"
#include "imex_kernels.h"

using namespace imex;

struct imex_handle {
  imex_config cfg;
  Params P;
  Grid G;
  hipStream_t stream = nullptr;
  int device = 0;
  bool fp32 = false;  // the FP32 form: float state, right-hand side and fluxes
  int n_cu = 256;  // compute units of the device (persistent grid sizes)
  int ntiles = 0;
  int par = 0;  // slot of the active-tile list filled by the latest k_mask
  unsigned int mask_seq = 0;  // k_mask launches so far (candidate-list slots)
  bool mask_full_next = true;  // the next k_mask visits every tile
  size_t lds = 0;
  std::vector<void*> allocs;
  std::vector<size_t> alloc_sizes;  // bytes of each entry of allocs (0 once freed)
  double* d_tmp = nullptr;  // N*nv scratch for layout conversion
  double* d_diag = nullptr;
  int diag_blocks = 0;
  double* dts_buf = nullptr;  // device copy of the per-step dt of a run_steps call
  int dts_cap = 0;
  // outputs (imex_output.hip)
  unsigned char* hpos_old = nullptr;  // [N] runout mask stored by the previous call
  void* d_red = nullptr;              // per-block partials of the runout reductions
  int red_cap = 0;
  // snapshots of output_solution's records (imex_output.hip): device staging
  // per record kind, copied to the host on their own stream
  bool bcent_set = false;  // imex_set_topography was called
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_staged = nullptr, ev_copied = nullptr;
  bool copy_inflight = false;
  double* stage[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t stage_cap[7] = {0, 0, 0, 0, 0, 0, 0};
  // pore-pressure runs (imex_pore.hip): primitive states, face states,
  // face fluxes and wave-speed maxima of the whole grid
  double *pp_qp = nullptr, *pp_qpL = nullptr, *pp_qpR = nullptr, *pp_qpB = nullptr,
         *pp_qpT = nullptr, *pp_Hx = nullptr, *pp_Hy = nullptr, *pp_amx = nullptr,
         *pp_amy = nullptr;
  std::string err;
};

// releases what imex_output.hip created (streams, events); called by imex_destroy
void imex_output_release(imex_handle* h);

// pore-pressure runs (imex_pore.hip): the buffers of the general path, and
// the launches of one device-driven step after its active-set pass
int imex_pore_alloc(imex_handle* h);
int imex_pore_step(imex_handle* h, int step_idx);

// the device-side driver's dt: the CFL step clipped to t_end and limited to
// 1.1 * (dt_old + dt_old_old) / 2, src/IMEX_SfloW2D.f90:417-439
__device__ __forceinline__ double driver_dt(const imex::Params& P, const imex::DevState* st) {
  double dt = imex::fmin_f(P.max_dt, __longlong_as_double((long long)st->dtmin_bits));
  const double t = st->t;
  if (t + dt > st->t_end) dt = st->t_end - t;
  return imex::fmin_f(dt, 1.1 * 0.5 * (st->dt_old + st->dt_old_old));
}

static inline int set_err(imex_handle* h, const char* what, hipError_t e) {
  if (h) {
    h->err = std::string(what) + ": " + hipGetErrorString(e);
  }
  return -1;
}
#define HCHK(h, call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return set_err((h), #call, e_); \
  } while (0)

template <typename T>
static inline int dalloc(imex_handle* h, T** p, size_t n) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, n ? n * sizeof(T) : 16);
  if (e != hipSuccess) return set_err(h, "hipMalloc", e);
  HCHK(h, hipMemsetAsync(v, 0, n ? n * sizeof(T) : 16, h->stream));
  h->allocs.push_back(v);
  h->alloc_sizes.push_back(n ? n * sizeof(T) : 16);
  *p = (T*)v;
  return 0;
}

// frees one dalloc allocation (its user must be done with it: the caller
// waits first) and drops it from the handle's list and footprint
static inline int dfree(imex_handle* h, void* p) {
  if (!p) return 0;
  for (size_t i = 0; i < h->allocs.size(); i++)
    if (h->allocs[i] == p) {
      h->allocs[i] = nullptr;
      h->alloc_sizes[i] = 0;
    }
  HCHK(h, hipFree(p));
  return 0;
}

