// Internal types of the batched column ID (shared by k_batched.hip and the host).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// Limits of the persistent kernel: the residual norms and the sampler's
// weights of a problem live in LDS (n <= RBRP_BATCH_MAXN) and the panel of a
// block is factorised by one wave (b <= RBRP_BATCH_MAXB).
#define RBRP_BATCH_MAXN 1024
#define RBRP_BATCH_MAXB 32

// Device copy of one problem (the public rbrp_problem_t minus host-only fields).
struct RbrpBatchProblem {
  const double* A;
  int64_t m, n, lda;
  uint64_t seed, stream_offset;
  int32_t* perm;
  double* X;
  int64_t ldx;
};

struct RbrpBatchResult {
  int64_t k;
  int64_t nblocks;
  double err_final;
  unsigned long long draws;
  int32_t status;
  int32_t pad;
};

// Launch shape: workspace slots (one per workgroup) and loop options.
struct RbrpBatchShape {
  void* ws;           // grid x slot_bytes
  size_t slot_bytes;  // Aw | R | Qbuf | Cbuf | Tbuf
  size_t off_R, off_Q, off_C, off_T;
  int64_t ldr;        // leading dimension of R, Cbuf, Tbuf (= largest kmax)
  int n_max;          // largest n (LDS arrays)
  int b;
  int filter;
  int max_rank;
  double tol2;
  int tile;  // row tiles a wave takes at once in the fused projection + update (1, 2 or 4)
};

size_t rbrp_batched_lds_bytes(int n_max, int b);

// Sets the kernel's dynamic-LDS limit for (n_max, b); call before launching
// (and before recording a launch into a graph).
hipError_t rbrp_batched_prepare(int n_max, int b);

// Workgroups of the persistent kernel (the instantiation for `tile`) that fit
// on one compute unit at once (registers and LDS for (n_max, b)); call after
// rbrp_batched_prepare.
int rbrp_batched_max_blocks_per_cu(int n_max, int b, int tile);

// Resets the queue counter and launches the persistent kernel on `grid` workgroups.
hipError_t rbrp_launch_colid_batched(hipStream_t st, int grid, const RbrpBatchProblem* probs,
                                     const int* order, int nprob, int* queue,
                                     RbrpBatchResult* results, const RbrpBatchShape& shape);
