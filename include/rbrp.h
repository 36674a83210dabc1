/*
 * rbrp.h -- C ABI of the MI355X-native RBRP column interpolative decomposition.
 *
 * Plain C: pointers, 64-bit sizes, status codes.  No torch or HIP types appear
 * in the signatures (streams are passed as `void*` holding a hipStream_t).
 *
 * Conventions shared by every entry point:
 *   - matrices are column-major FP64 device pointers with a leading dimension;
 *   - the input matrix is never written;
 *   - return value: 0 = success, -i = argument i is invalid (LAPACK style,
 *     1-based), +e = HIP runtime error code e, RBRP_ERR_NUMERIC for a
 *     non-finite input or a breakdown; nothing throws, exits or asserts;
 *   - a handle owns its workspace and stream and no global state, so distinct
 *     handles may be used concurrently from different host threads.
 */
#ifndef RBRP_H_
#define RBRP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBRP_ERR_NUMERIC 100000
#define RBRP_ERR_ALLOC 100001
/* The problem's workspace does not fit the static scratch of rbrp_colid_d_noalloc. */
#define RBRP_ERR_SCRATCH 100002

/* Size of the per-device static scratch array used by rbrp_colid_d_noalloc. */
#define RBRP_NOALLOC_SCRATCH_BYTES (64ull << 20)

typedef struct rbrp_handle_s* rbrp_handle_t;

/* Filter threshold (gpu/rbrp.cu:198-207 vs RBRP.ipynb c11:69-73). */
enum { RBRP_FILTER_SAMPLED = 0, RBRP_FILTER_NOMINAL = 1 };
/*
 * Arithmetic of the residual.  FP64 is the reference's.  FP32 is the mixed
 * form: the trailing residual is stored in FP32 and projected / updated with
 * FP32 MFMA products and FP32 accumulation (split-K partials and column norms
 * summed in FP64), while the sampling weights' sum and stopping test, the
 * re-orthogonalisation, the pivoted factorisation of the sampled block, R and
 * X stay FP64.  Its resolution is bounded by FP32 storage: relative errors
 * below about 1e-7 ||A||_F are not resolved.
 */
enum { RBRP_PREC_FP64 = 0, RBRP_PREC_FP32 = 1 };

typedef struct {
  int64_t b;              /* nominal block size, 1..128 (rbrp.hpp:6 `blk`) */
  double tol;             /* stop when ||A - A(:,J)X||_F <= tol ||A||_F; 0 = run to min(m,n) */
  int64_t max_rank;       /* 0 = min(m,n) */
  uint64_t seed;          /* Philox4x64-10 key word 0 (word 1 is 0) */
  uint64_t stream_offset; /* Philox initial counter (numpy.random.Philox `counter`) */
  int32_t precision;      /* RBRP_PREC_* */
  int32_t filter;         /* RBRP_FILTER_* */
} rbrp_opts_t;

typedef struct {
  /* caller-provided outputs */
  int32_t* perm;       /* device, n entries: perm[0:k] = skeleton J, perm[k:n] = redundant */
  double* X;           /* device, X = R11^{-1} R12 (k x (n-k)), column-major, or NULL to skip */
  int64_t ldx;         /* leading dimension of X, >= min(m,n) (or >= max_rank) */
  int64_t trace_cap;   /* capacity of the three host trace arrays below (0 = none) */
  int64_t* ranks;      /* host: cumulative rank after each block          (rbrp.cu:475-482) */
  double* error;       /* host: squared relative residual after each block (rbrp.cu:467) */
  double* t_ranks;     /* host: cumulative seconds after each block       (rbrp.cu:477-483) */
  /* filled by the call */
  int64_t k;           /* rank found */
  int64_t nblocks;     /* number of blocks run */
  double t_total;      /* seconds from entry to return (rbrp.cu:485) */
  double err_final;    /* non-squared final relative residual */
  uint64_t draws;      /* uniforms consumed from the Philox stream */
} rbrp_result_t;

/* Per-phase device time (hipEvents), launches and algorithmic work. */
enum {
  RBRP_PHASE_NORMS = 0,    /* copy + squared column norms of A (one kernel) */
  RBRP_PHASE_SAMPLE,       /* weighted sampling without replacement (one kernel) */
  RBRP_PHASE_CGS,          /* re-orthogonalisation of the sampled block */
  RBRP_PHASE_PANEL,        /* Gram, pivoted Cholesky, filter, CholQR2 of the block */
  RBRP_PHASE_PERMUTE,      /* symmetric-difference column permutation */
  RBRP_PHASE_PROJECT,      /* W = Q_b^T A_rem: the split-K MFMA GEMM kernel alone */
  RBRP_PHASE_PROJECT_RED,  /* reduction of the split-K partials into R */
  RBRP_PHASE_UPDATE,       /* A_rem -= Q_b W with fused norms: the NN MFMA kernel alone */
  RBRP_PHASE_RESID,        /* residual norms / sum from the update's partials */
  RBRP_PHASE_SOLVE,        /* X = R11^{-1} R12 */
  RBRP_PHASE_COUNT
};

typedef struct {
  double ms[RBRP_PHASE_COUNT];      /* summed device milliseconds */
  int64_t launches[RBRP_PHASE_COUNT];
  double flops[RBRP_PHASE_COUNT];   /* algorithmic flops */
  double bytes[RBRP_PHASE_COUNT];   /* algorithmic HBM bytes */
} rbrp_phase_stats_t;

/* Handle lifetime.  `stream` is a hipStream_t (NULL = the default stream). */
int rbrp_create(rbrp_handle_t* handle, void* stream);
int rbrp_destroy(rbrp_handle_t handle);
int rbrp_set_stream(rbrp_handle_t handle, void* stream);
/* Per-phase hipEvent timing: 0 off, 1 every phase, 2 only the projection and
 * update phases (RBRP_PHASE_PROJECT, RBRP_PHASE_UPDATE: two events per block
 * around each, the least added to the stream).  reset clears the stats. */
int rbrp_set_timing(rbrp_handle_t handle, int enable);
int rbrp_get_phase_stats(rbrp_handle_t handle, rbrp_phase_stats_t* out);
int rbrp_reset_phase_stats(rbrp_handle_t handle);
/* Release the cached workspace. */
int rbrp_release_workspace(rbrp_handle_t handle);

/*
 * FP64 RBRP column ID.  Replaces `RBRP(int m, int n, double *A, int blk,
 * std::vector<int>& ranks, std::vector<double>& error, std::vector<double>& t_ranks,
 * double& t_rbrp, double& flops)` (gpu/rbrp.hpp:6-8, gpu/rbrp.cu:333-503) and adds
 * the tolerance, skeleton/permutation and X outputs of matlab/RBRP.m:2 and
 * matlab/driver.m:83-94.  dA is m x n, leading dimension lda >= m.
 */
int rbrp_colid_d(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                 const rbrp_opts_t* opts, rbrp_result_t* result);

/*
 * rbrp_colid_d without any device allocation.  No counterpart in the reference
 * (gpu/rbrp.cu:333-503 allocates per call); it is for callers that cannot
 * allocate: code running while a HIP graph is being captured, tight serving
 * loops, harnesses that forbid hipMalloc.  Same arguments, results and status
 * codes as rbrp_colid_d; the workspace comes from a static device array of
 * RBRP_NOALLOC_SCRATCH_BYTES per GPU (held by each GPU the library runs on
 * until the process ends), and a problem whose workspace does not fit it
 * (roughly m*n + k*n doubles plus small buffers: e.g. 2000 x 1500) is rejected
 * with RBRP_ERR_SCRATCH before anything is launched.  The scratch address is
 * looked up by rbrp_create on the handle's device.  Phase timing is not
 * recorded by this entry point (timing needs events).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other handles,
 * streams or threads) race and must be serialised by the caller.  When the
 * arena form's split-K partials fit, the launches and results are identical
 * to rbrp_colid_d's.
 */
int rbrp_colid_d_noalloc(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                         const rbrp_opts_t* opts, rbrp_result_t* result);

/*
 * The same column ID for an FP32-stored input (dA is m x n FP32, leading
 * dimension lda).  With opts->precision = RBRP_PREC_FP32 the residual stays
 * FP32; with RBRP_PREC_FP64 the loop runs on an FP64 copy.  No reference
 * counterpart: the reference takes FP64 input only (gpu/rbrp.hpp:6-8).
 */
int rbrp_colid_s(rbrp_handle_t handle, const float* dA, int64_t m, int64_t n, int64_t lda,
                 const rbrp_opts_t* opts, rbrp_result_t* result);

/*
 * Gaussian sketch: G(row, col) = scale * z_t, t = row + col * rows, z from the
 * Box-Muller pairs of the Philox stream with key (seed, 1) and counter
 * `offset`.  Replaces Random::Gaussian (gpu/random.cu:44-48).  dG is rows x cols
 * (device, leading dimension ldg >= rows).
 */
int rbrp_gaussian_sketch_d(rbrp_handle_t handle, uint64_t seed, uint64_t offset, int64_t rows,
                           int64_t cols, double scale, double* dG, int64_t ldg);

/*
 * Randomised CPQR ID with 2x oversampling.  Replaces `void RandCPQR_OS(const
 * double *A, int m, int n, int rank, int *h_sk, int *h_rd, double *&T, double&
 * flops)` (gpu/rid.hpp:19-20, gpu/randCPQR_OS.cu:20-111): Y = G A with G
 * (2r x m) ~ N(0, 1/m); pivoted QR of Y(1:r, :) gives sk = jpvt[0:r] and
 * rd = jpvt[r:n] (host arrays, as in the reference); T (r x (n-r), device,
 * leading dimension ldt, caller-allocated) by the oversampled interpolation.
 * dG (2r x m, ldg) is an optional caller sketch; NULL draws it from (seed, offset).
 */
int rbrp_randcpqr_os_d(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                       int64_t rank, uint64_t seed, uint64_t offset, const double* dG, int64_t ldg,
                       int32_t* h_sk, int32_t* h_rd, double* dT, int64_t ldt);

/*
 * Randomised LUPP ID with 2x oversampling.  Replaces `void RandLUPP_OS(const
 * double *A, int m, int n, int rank, int *&sk, int *&rd, double *&T, double&
 * flops)` (gpu/rid.hpp:12-13, gpu/randLUPP_OS.cu:14-119): Y = G A (G is 2k x m);
 * LU with partial pivoting of Y(1:k, :)^T; the row interchanges applied to the
 * identity give sk = P[0:k], rd = P[k:n] (device arrays here, caller-allocated);
 * T as for rbrp_randcpqr_os_d.
 */
int rbrp_randlupp_os_d(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                       int64_t rank, uint64_t seed, uint64_t offset, const double* dG, int64_t ldg,
                       int32_t* d_sk, int32_t* d_rd, double* dT, int64_t ldt);

/*
 * Full-matrix CPQR ID baseline.  Replaces `void CPQR(int m, int n, double *A,
 * const std::vector<int> ranks, std::vector<double> &err, std::vector<double>
 * &t_ranks, double &t_cpqr, double &flops)` (gpu/id_cpqr.hpp:6-8,
 * gpu/id_cpqr.cu:30-172): Householder QR with column pivoting (first maximum
 * of the trailing squared column norms, dgeqp3's pivot rule) of a copy of A
 * (A is not written) to min(m, n) steps, timed into *t_cpqr (seconds).  For
 * each ranks[t] (1 <= r <= min(m, n)): T = R11^{-1} R12 and error[t] =
 * ||A(:, P(r:n)) - A(:, P(0:r)) T||_F^2 / ||A||_F^2 (squared, as the
 * reference); t_ranks[t] = the partial-factorization time model of
 * id_cpqr.cu:93-104 plus the solve's time.  *flops = FLOPS_DGEQRF(m, n) +
 * sum_t r^2 (n - r).  error and t_ranks are host arrays of nranks; d_perm
 * (optional, device, n entries) receives the pivot order (0-based).
 */
int rbrp_cpqr_id_d(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                   int64_t nranks, const int32_t* ranks, double* error, double* t_ranks,
                   double* t_cpqr, double* flops, int32_t* d_perm);

/*
 * Test matrices of the reference's driver, generated on the host (no GPU
 * needed) into A (column-major, leading dimension = its row count) from a
 * seeded std::mt19937 stream.  Replace Matrix_GE, Matrix_SNN and Matrix_GMM
 * (gpu/driver.cpp:23-47, :50-91, :93-111), which draw from unseeded
 * generators: same families, not the same matrices.
 *   GE:  n x n, U diag(s) V^T with U, V the Q factors of uniform [-1, 1]
 *        matrices, s_i = 1 (i < 100), max(0.8^(i-100), 1e-5) after.
 *   SNN: n x n, (U D) V with sparse U, V (ceil(s n) random entries per
 *        column on average, uniform [0, 1)), D = diag(0.1 (i < 100), 1/(i-99)).
 *   GMM: d x n, N(0, 1) plus 10 (j + 1) on row j of the n/100 columns of
 *        cluster j (100 clusters); needs n >= 100 and d > (n - 1) / (n / 100).
 */
int rbrp_matrix_ge_h(int64_t n, uint64_t seed, double* A);
int rbrp_matrix_snn_h(int64_t n, double s, uint64_t seed, double* A);
int rbrp_matrix_gmm_h(int64_t n, int64_t d, uint64_t seed, double* A);

/*
 * Randomised CPQR ID without oversampling.  Replaces `void RandCPQR(const
 * double *A, int m, int n, int rank, int *h_sk, int *h_rd, double *&T, double&
 * flops)` (gpu/rid.hpp:16-17, gpu/randCPQR.cu:15-67): Y = G A with G (r x m) ~
 * N(0, 1/m); pivoted QR of all of Y gives sk = jpvt[0:r], rd = jpvt[r:n] (host
 * arrays); T = R11^{-1} R12 from the same factor (r x (n-r), device, ldt).
 * dG (r x m, ldg) is an optional caller sketch; NULL draws it from (seed, offset).
 */
int rbrp_randcpqr_d(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                    int64_t rank, uint64_t seed, uint64_t offset, const double* dG, int64_t ldg,
                    int32_t* h_sk, int32_t* h_rd, double* dT, int64_t ldt);

/*
 * Randomised LUPP ID without oversampling.  Replaces `void RandLUPP(const
 * double *A, int m, int n, int rank, int *&sk, int *&rd, double *&T, double&
 * flops)` (gpu/rid.hpp:9-10, gpu/randLUPP.cu:13-107): Y = G A (G is k x m); LU
 * with partial pivoting of Y^T (n x k) = P^T [L11; L21] U; sk = P[0:k],
 * rd = P[k:n] (device arrays, caller-allocated); T = (L21 L11^{-1})^T.
 */
int rbrp_randlupp_d(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                    int64_t rank, uint64_t seed, uint64_t offset, const double* dG, int64_t ldg,
                    int32_t* d_sk, int32_t* d_rd, double* dT, int64_t ldt);

/*
 * Adaptive randomised LUPP ID.  Replaces `void RandAdapLUPP(const double *A,
 * int m, int n, int *&sk, int *&rd, double *&T, int &rank, double &flops,
 * double tol=1e-8, int blk=16)` (gpu/rid.hpp:5-7, declared there and never
 * defined under gpu/), with the semantics of blocked_rlupp_id
 * (python/cache.ipynb c8:80-143) in column form.  Block t draws a Gaussian
 * sketch Y_t = G_t A (blk rows, G_t ~ N(0, 1/blk) from the Philox stream with
 * key (seed, 1) and counter offset + t*blk*m).  The residual of Y_0 is Y_0;
 * each block takes an LU with partial pivoting of the candidates' residual
 * transposed (candidates x blk): its first blk pivots join the skeleton, the
 * interpolation matrix M = L2 L1^{-1} of the accumulated unit-lower factor is
 * extended, and the next sketch gives both the residual
 * Y(:, cand)^T - M Y(:, skel)^T of the next block and the error estimate
 * ||that residual||_F (an unbiased estimate of ||A(:, rd) - A(:, sk) T||_F).
 * Stops when the estimate is at most tol * ||A||_F (relative; the notebook
 * compares it with an absolute tol) or the rank reaches min(m, n).
 * Outputs (device, caller-allocated with worst-case sizes): d_sk (n entries;
 * the first *rank are the skeleton), d_rd (n entries; the first n - *rank are
 * the rest), dT (*rank x (n - *rank), leading dimension ldt >= min(m, n)),
 * a block being cut short where it would take the rank past min(m, n),
 * T = M^T so that A(:, rd) ~ A(:, sk) T.  err_est (optional) receives the
 * last estimate relative to ||A||_F, flops (optional) the arithmetic count.
 * blk is 1..32.
 */
int rbrp_randadaplupp_d(rbrp_handle_t handle, const double* dA, int64_t m, int64_t n, int64_t lda,
                        double tol, int64_t blk, uint64_t seed, uint64_t offset, int32_t* d_sk,
                        int32_t* d_rd, double* dT, int64_t ldt, int64_t* rank, double* err_est,
                        double* flops);

/*
 * Oversampled interpolation matrix.  Replaces Compute_interpolation_matrix
 * (gpu/interp.cu:45-105): Yp = Y(:, P); QR of Yp(:, 0:k); T = R^{-1} Q^T Yp(:, k:n)
 * (k x (n-k), leading dimension ldt).  Y is l x n (ldy); P (device) has n entries.
 */
int rbrp_interp_os_d(rbrp_handle_t handle, const double* dY, int64_t l, int64_t n, int64_t ldy,
                     int64_t k, const int32_t* dP, double* dT, int64_t ldt);

/*
 * Batched column IDs (many independent FP64 problems, e.g. the off-diagonal
 * blocks of a hierarchical kernel matrix).  No reference counterpart: the
 * reference runs one matrix per call (gpu/rbrp.hpp:6-8, gpu/driver.cpp:176).
 * Every problem follows rbrp_colid_d with the shared options (b, tol,
 * max_rank, filter; opts->seed and opts->stream_offset are replaced by the
 * problem's own), so a problem gives the same result in any batch.
 */
typedef struct {
  const double* A;        /* device, m x n column-major, leading dimension lda */
  int64_t m, n, lda;
  uint64_t seed;          /* Philox key word 0 of this problem's stream */
  uint64_t stream_offset; /* its initial counter */
  int32_t* perm;          /* device, n entries: perm[0:k] skeleton, perm[k:n] redundant */
  double* X;              /* device, k x (n-k) with leading dimension ldx, or NULL */
  int64_t ldx;            /* >= min(m, n) (>= max_rank when set) when X is given */
} rbrp_problem_t;

typedef struct {
  int64_t k;          /* rank found */
  int64_t nblocks;    /* blocks run */
  double err_final;   /* non-squared final relative residual */
  uint64_t draws;     /* uniforms consumed */
  int32_t status;     /* 0 ok; nonzero: numerical failure of this problem (outputs undefined) */
  int32_t pad;
} rbrp_problem_result_t;

/* RBRP_BATCH_PERSISTENT: one persistent launch in which each workgroup takes
 * whole problems from a device work queue (largest first) and runs their
 * complete loop; the launch is recorded in a hipGraph and replayed by later
 * calls with the same launch shape.  Needs n <= 1024 and b <= 32 for every
 * problem.  RBRP_BATCH_LOOP: one rbrp_colid_d call per problem.
 * RBRP_BATCH_AUTO: persistent when the limits allow, else the loop. */
enum { RBRP_BATCH_AUTO = 0, RBRP_BATCH_PERSISTENT = 1, RBRP_BATCH_LOOP = 2 };

/* results: host array of nprob entries.  Returns 0, -i for a bad argument i,
 * or a HIP error; per-problem numerical failures are reported in results. */
int rbrp_colid_batched_d(rbrp_handle_t handle, int64_t nprob, const rbrp_problem_t* problems,
                         const rbrp_opts_t* opts, rbrp_problem_result_t* results, int mode);

/* Autotuner of the persistent launch: runs the batch once per candidate
 * launch shape -- the row tiles (1, 2 or 4) a wave takes at once in the fused
 * projection + trailing update, which sets the kernel's register footprint,
 * times the number of workgroups per compute unit -- keeps the fastest for
 * later calls of the same batch shape on this handle and returns its
 * workgroups per CU in *wg_per_cu (results of the last run are left in the
 * problems' outputs).  b is not tuned: it changes the result.  No counterpart
 * in the reference: the north star's "tile shapes chosen by an autotuner". */
int rbrp_batched_tune(rbrp_handle_t handle, int64_t nprob, const rbrp_problem_t* problems,
                      const rbrp_opts_t* opts, int* wg_per_cu);

/* The launch shape later persistent calls for this batch shape will use on
 * this handle: the tuner's pick, else 4 workgroups per CU and 2 row tiles.
 * No counterpart in the reference (reports the autotuner's choice). */
int rbrp_batched_launch_shape(rbrp_handle_t handle, int64_t nprob, const rbrp_problem_t* problems,
                              const rbrp_opts_t* opts, int* wg_per_cu, int* tile_rows);

/* Defaults: b = 64, tol = 0, max_rank = 0, seed = 0, offset = 0, FP64, sampled filter. */
void rbrp_opts_default(rbrp_opts_t* opts);

/*
 * Host-side Philox4x64-10 block (no GPU needed): out[0..3] = Philox(ctr, key).
 * Exposes the generator the device sampler uses so that it can be checked
 * against its known-answer vectors.
 */
void rbrp_philox4x64_10_host(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* Fill out[0..count) with the uniforms draw0, draw0+1, ... of the stream
 * (seed, offset), computed on the device.  dOut is a device pointer. */
int rbrp_uniforms_d(rbrp_handle_t handle, uint64_t seed, uint64_t offset, uint64_t draw0,
                    int64_t count, double* dOut);

/* Library version string. */
const char* rbrp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RBRP_H_ */
