// Internal definition of the library handle (one stream + cached workspaces).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/rbrp.h"
#include "rbrp_common.h"

struct RbrpWorkspace {
  void* arena = nullptr;
  size_t arena_bytes = 0;
  // carved pointers of the RBRP loop
  double *Aw, *R, *diags, *w, *cdf, *G, *G2, *Rb, *Rinv, *R2, *R2inv, *Cbuf, *Qbuf, *P, *normpart,
      *Wt;
  double* Dinv;  // diagonal-block inverses of R for the X solve
  // mixed precision: FP32 working matrix and row-major W, FP64 skeleton basis
  // and FP64 copy of the sampled block (null in the FP64 loop, where Aw, Wt
  // are used instead)
  float *Aw32, *Wt32;
  double *Q64, *Vd;
  // mixed loop: chunk totals and total of the FP64 residual recompute
  double *rcol, *r64;
  int *J, *Jk, *Jin, *Jout, *Jkpos;
  RbrpBlockState* st;
  RbrpBlockState* st_snap;  // the last published block record (k_resid_total)
  int64_t Pcap;
};

struct rbrp_handle_s {
  hipStream_t stream = nullptr;
  int device = 0;
  // workgroups of the split-K projection kernels resident at once (0: unknown)
  int tn64_resident = 0, tn32_resident = 0, tn32w_resident = 0, nn64_resident = 0;
  RbrpWorkspace ws;
  void* sk_arena = nullptr;  // workspace of the sketch comparators
  size_t sk_bytes = 0;
  RbrpBlockState* host_state = nullptr;  // pinned, host-coherent
  int* host_scalar = nullptr;             // pinned: small read-backs (the QRCP panel width)
  int pub_seq = 0;                        // last publication number handed out
  int timing = 0;  // 0 off, 1 every phase, 2 projection and update only
  rbrp_phase_stats_t stats;
  std::vector<hipEvent_t> events;
  // batched calls: slot workspace, problem metadata, pinned staging, a
  // private stream for graph capture, the recorded launch and the tuner's pick
  void* bt_ws = nullptr;
  size_t bt_ws_bytes = 0;
  void* bt_meta = nullptr;
  size_t bt_meta_bytes = 0;
  void* bt_host = nullptr;
  size_t bt_host_bytes = 0;
  hipStream_t bt_stream = nullptr;
  hipGraph_t bt_graph = nullptr;
  hipGraphExec_t bt_exec = nullptr;
  std::vector<uint64_t> bt_key;
  // shape key -> the tuner's launch shape, workgroups per CU | update row tiles << 8
  std::vector<std::pair<std::vector<uint64_t>, int>> bt_tuned;
};

// Releases the batched-call resources of a handle (rbrp_batched.cpp).
void rbrp_batched_release(rbrp_handle_t h);

// Grow-only device buffer of the handle for the sketch comparators.
hipError_t rbrp_sketch_arena(rbrp_handle_t h, size_t bytes, void** out);

// X (k x nc, ldx) <- R^{-1} X, R upper triangular (k x k, ldr); blocked, on st.
// dinv (optional): scratch of 4 * ceil(k / 256) * 4096 doubles for the
// diagonal-block inverses.
hipError_t rbrp_trsm_upper_left(hipStream_t st, int64_t k, int64_t nc, double* R, int64_t ldr,
                                double* X, int64_t ldx, int* launches, double* dinv = nullptr);
