/*
 * rbrp_testing.h -- kernel-level entry points used by the parity tests.
 *
 * These have no counterpart in the reference: they expose single kernels of
 * the RBRP library (the MFMA GEMMs and the panel factorisation) so that each
 * can be checked against a host computation in isolation.  Same conventions as
 * rbrp.h: column-major FP64 device pointers, status codes, `stream` is a
 * hipStream_t passed as void*.  Every call synchronises its stream.
 */
#ifndef RBRP_TESTING_H_
#define RBRP_TESTING_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C (M x N, ldc) = A^T B with A (K x M, lda), B (K x N, ldb) by the split-K
 * MFMA kernel and its reduction.  K must be a multiple of 16. */
int rbrp_test_gemm_tn(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int nsplit);

/* C (M x N) = A B (mode 0), C -= A B (mode 1) or C += A B (mode 2) with A
 * (M x K, lda) and B (K x N, ldb) by the NN MFMA kernel.  If normpart is not
 * NULL it receives ceil(M/64) x N per-tile squared column norms of the new C
 * (leading dimension N). */
int rbrp_test_gemm_nn(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int mode,
                      double* normpart);

/* Trailing-update kernel (K <= 128): C (M x N) -= A Bt^T-layout product, i.e.
 * C(r, c) -= sum_k A(r, k) Bt[k * ldbt + c] with A (M x K, lda) column-major
 * and Bt row-major.  normpart (optional, ceil(M/64) x N, leading dimension N,
 * zero-initialised by the caller) receives partial squared column norms of the
 * new C over row groups: its rows sum to the column norms. */
int rbrp_test_gemm_upd(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                       const double* Bt, int64_t ldbt, double* C, int64_t ldc, double* normpart);

/* FP64 projection kernel (64 x 128 tiles): C (M x N, ldc) = A^T B with A
 * (K x M) and B (K x N) column-major with the common leading dimension ld (even),
 * K a multiple of 16, split over about nsplit chunks of K. */
int rbrp_test_gemm_tn64(void* stream, int64_t M, int64_t N, int64_t K, const double* A,
                        const double* B, int64_t ld, double* C, int64_t ldc, int nsplit);

/* Symmetric split-K Gram kernel of the panel: C (N x N, ldc) = A^T A with A
 * (K x N, lda even) column-major, N <= 128, K a multiple of 16, split over
 * about nsplit chunks of K. */
int rbrp_test_gram(void* stream, int64_t N, int64_t K, const double* A, int64_t lda, double* C,
                   int64_t ldc, int nsplit);

/* The blocked back substitution of the X solve: X (k x nc, ldx) <- R^{-1} X for
 * the k x k upper-triangular R (ldr) column-major on the device; with
 * use_dinv the diagonal-block inverses and the one-launch column-tile solve,
 * else the substitution path. */
int rbrp_test_trsm(void* stream, int64_t k, int64_t nc, const double* R, int64_t ldr, double* X,
                   int64_t ldx, int use_dinv);

/* Workgroups of the split-K projection kernel resident at once on the current
 * device: which = 0 for the FP64 kernel, 1 for the FP32 one (0 if unknown). */
int rbrp_test_resident(int which);

/* FP64 tall NN kernel (64 x 64 tiles, long reduction): C (M x N, ldc) (op)=
 * A B with A (M x K, lda even) and B (K x N, ldb) column-major; M a multiple
 * of 4; mode 0 sets, 1 subtracts, 2 adds. */
int rbrp_test_gemm_nn64(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double* C, int64_t ldc, int mode);

/* FP32 split-K TN kernel: C (M x N, FP64, ldc) = A^T B with A (K x M) and B
 * (K x N) FP32 column-major with the common leading dimension ld (a multiple
 * of 4), FP32 products and accumulation per split, FP64 split sum.  K must be
 * a multiple of 32. */
int rbrp_test_gemm_tn32(void* stream, int64_t M, int64_t N, int64_t K, const float* A,
                        const float* B, int64_t ld, double* C, int64_t ldc, int nsplit);

/* FP32 trailing-update kernel (K <= 128, M, lda, ldc even): C(r, c) -= sum_k
 * A(r, k) Bt[k * ldbt + c], FP32 data; normpart (optional, FP64, ceil(M/64)
 * x N, leading dimension N, zero-initialised by the caller) receives partial
 * squared column norms over row groups: its rows sum to the column norms. */
int rbrp_test_gemm_upd32(void* stream, int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                         const float* Bt, int64_t ldbt, float* C, int64_t ldc, double* normpart);

/* Pivoted Cholesky + robust filter of an s x s Gram matrix G (ldg = s) whose
 * columns correspond to the column indices J[0:s] (device), at block start i.
 * Outputs (device, capacity b each): Jk[0:bt] (pivot order), Rb (b x b, ld b),
 * Jin/Jout[0:nswap]; bt, nswap and status are written to host ints. */
int rbrp_test_panel(void* stream, const double* G, int64_t s, int64_t b, int filter, int64_t i,
                    const int32_t* J, int32_t* Jk, double* Rb, int32_t* Jin, int32_t* Jout,
                    int32_t* bt, int32_t* nswap, int32_t* status);

/* The same panel step with the kept factor's inverse: Rinv (device, b x b,
 * ld b) receives Rb(0:bt, 0:bt)^{-1}, zero below the diagonal. */
int rbrp_test_panel_inv(void* stream, const double* G, int64_t s, int64_t b, int filter,
                        int64_t i, const int32_t* J, int32_t* Jk, double* Rb, double* Rinv,
                        int32_t* Jin, int32_t* Jout, int32_t* bt, int32_t* nswap,
                        int32_t* status);

/* rbrp_test_panel_inv on a chosen kernel: 0 the library's choice for b, 1 the
 * one-wave kernel (b <= 64), 2 the register-tile workgroup kernel (b <= 128). */
int rbrp_test_panel_kernel(void* stream, const double* G, int64_t s, int64_t b, int filter,
                           int64_t i, const int32_t* J, int32_t* Jk, double* Rb, double* Rinv,
                           int32_t* Jin, int32_t* Jout, int32_t* bt, int32_t* nswap,
                           int32_t* status, int kernel);

/* Second Cholesky-QR pass of a block of bt kept columns (b >= bt the block
 * capacity): from G2 (device, bt x bt, ld bt) and the first factor Rb (device,
 * b x b, ld b) it writes R2 = chol(G2) and R2inv = R2^{-1} (device, b x b,
 * ld b) and Rout = R2 * Rb (device, bt x bt, ld bt); status 2 reports a
 * non-positive pivot. */
int rbrp_test_chol2(void* stream, const double* G2, int64_t bt, int64_t b, const double* Rb,
                    double* R2, double* R2inv, double* Rout, int32_t* status);

/* rbrp_test_chol2 on a chosen kernel (0, 1, 2 as for rbrp_test_panel_kernel). */
int rbrp_test_chol2_kernel(void* stream, const double* G2, int64_t bt, int64_t b, const double* Rb,
                           double* R2, double* R2inv, double* Rout, int32_t* status, int kernel);

/* Sampler kernel alone: s = min(b, kmax - i, nnz(diags[i:n])) indices (absolute,
 * first-occurrence order) into J (device, capacity b); draws_in is the number
 * of uniforms already consumed; the new count and s go to host ints. */
int rbrp_test_sample(void* stream, const double* diags, int64_t i, int64_t n, int64_t kmax,
                     int64_t b, uint64_t seed, uint64_t offset, uint64_t draws_in, int32_t* J,
                     int32_t* s_out, uint64_t* draws_out);

/* Device time of the sampler per call over diags[0:n) (b draws per block,
 * Philox key (seed, 0)): one untimed call, then `reps` calls between two
 * events; *ms_per_call receives the mean.  Candidate ranges above 65536 take
 * the multi-workgroup form. */
int rbrp_test_sample_time(void* stream, const double* diags, int64_t n, int64_t b, uint64_t seed,
                          int reps, double* ms_per_call);

/* Ceilings of the part for the benchmark: kind 0 a streaming read of `bytes`
 * (rate in GB/s), 1 a streaming copy of `bytes` (GB/s read + written), 2 / 3
 * loops of independent FP64 / FP32 16x16x4 matrix-core instructions on every
 * compute unit (TFLOP/s), 4 an in-place read-modify-write sweep of a 50048-row
 * column-major matrix of about `bytes` in the trailing update's tile order
 * (GB/s read + written); two untimed launches, then `reps` between events. */
int rbrp_test_ceiling(void* stream, int kind, int64_t bytes, int reps, double* ms_per_launch,
                      double* rate);

#ifdef __cplusplus
}
#endif

#endif /* RBRP_TESTING_H_ */
