// Kernel-level entry points for the parity tests (include/rbrp_testing.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/rbrp_testing.h"
#include "rbrp_handle.h"
#include "rbrp_kernels.h"

#define HIPCHK(x)                         \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) return (int)e_; \
  } while (0)

extern "C" {

int rbrp_test_gemm_tn(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int nsplit) {
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (K < 0 || K % 16) return -4;
  if (nsplit < 1) return -11;
  if (M == 0 || N == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int64_t kc = (K + nsplit - 1) / nsplit;
  kc = std::max<int64_t>(16, (kc + 15) / 16 * 16);
  const int ns = (int)std::max<int64_t>(1, (K + kc - 1) / kc);
  double* P = nullptr;
  HIPCHK(hipMalloc(&P, sizeof(double) * M * N * ns));
  hipError_t e = rbrp_launch_gemm_tn(st, dim_static(M), dim_static(N), M, N, K, ns, kc,
                                     mat((double*)A, lda), mat((double*)B, ldb), P, M, M * N);
  if (e == hipSuccess)
    e = rbrp_launch_reduce_splits(st, dim_static(M), dim_static(N), M, N, P, M, M * N, ns,
                                  mat(C, ldc), GEMM_SET);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(P);
  return (int)e;
}

int rbrp_test_gemm_nn(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int mode,
                      double* normpart) {
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (K < 0) return -4;
  if (mode < 0 || mode > 2) return -11;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(rbrp_launch_gemm_nn(st, dim_static(M), M, dim_static(N), N, dim_static(K), mat((double*)A, lda),
                             mat((double*)B, ldb), mat(C, ldc), mode, normpart, N));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int rbrp_test_gemm_upd(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                       const double* Bt, int64_t ldbt, double* C, int64_t ldc, double* normpart) {
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (K < 0 || K > 128) return -4;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(rbrp_launch_gemm_trail(st, M, dim_static(N), N, dim_static(K), K, A, lda, 0, nullptr, Bt,
                                ldbt, C, ldc, 0, nullptr, normpart, N));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int rbrp_test_gemm_tn64(void* stream, int64_t M, int64_t N, int64_t K, const double* A,
                        const double* B, int64_t ld, double* C, int64_t ldc, int nsplit) {
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (K < 0 || K % 16) return -4;
  if (ld < std::max<int64_t>(K, 1) || ld % 2) return -7;
  if (nsplit < 1) return -10;
  if (M == 0 || N == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int64_t kc = (K + nsplit - 1) / nsplit;
  kc = std::max<int64_t>(16, (kc + 15) / 16 * 16);
  const int ns = (int)std::max<int64_t>(1, (K + kc - 1) / kc);
  double* P = nullptr;
  HIPCHK(hipMalloc(&P, sizeof(double) * M * N * ns));
  hipError_t e = rbrp_launch_gemm_tn64(st, dim_static(M), dim_static(N), M, N, K, ns, kc,
                                       mat((double*)A, ld), mat((double*)B, ld), P, M, M * N);
  if (e == hipSuccess)
    e = rbrp_launch_reduce_splits(st, dim_static(M), dim_static(N), M, N, P, M, M * N, ns,
                                  mat(C, ldc), GEMM_SET);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(P);
  return (int)e;
}

int rbrp_test_gram(void* stream, int64_t N, int64_t K, const double* A, int64_t lda, double* C,
                   int64_t ldc, int nsplit) {
  if (N < 0 || N > 128) return -2;
  if (K < 0 || K % 16) return -3;
  if (lda % 2 || lda < std::max<int64_t>(K, 1)) return -5;
  if (ldc < std::max<int64_t>(N, 1)) return -7;
  if (nsplit < 1) return -8;
  if (N == 0 || K == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int64_t kc = (K + nsplit - 1) / nsplit;
  kc = std::max<int64_t>(16, (kc + 15) / 16 * 16);
  const int ns = (int)std::max<int64_t>(1, (K + kc - 1) / kc);
  double* P = nullptr;
  HIPCHK(hipMalloc(&P, sizeof(double) * N * N * ns));
  hipError_t e = rbrp_launch_gram(st, dim_static(N), N, K, ns, kc, mat((double*)A, lda), P, N,
                                  N * N);
  if (e == hipSuccess)
    e = rbrp_launch_reduce_splits(st, dim_static(N), dim_static(N), N, N, P, N, N * N, ns,
                                  mat(C, ldc), GEMM_SET);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(P);
  return (int)e;
}

int rbrp_test_trsm(void* stream, int64_t k, int64_t nc, const double* R, int64_t ldr, double* X,
                   int64_t ldx, int use_dinv) {
  if (k < 0) return -2;
  if (nc < 0) return -3;
  if (ldr < std::max<int64_t>(k, 1)) return -5;
  if (ldx < std::max<int64_t>(k, 1)) return -7;
  if (k == 0 || nc == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  double* dinv = nullptr;
  if (use_dinv) HIPCHK(hipMalloc(&dinv, sizeof(double) * 4 * ((k + 255) / 256) * 4096));
  hipError_t e = rbrp_trsm_upper_left(st, k, nc, (double*)R, ldr, X, ldx, nullptr, dinv);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (dinv) (void)hipFree(dinv);
  return (int)e;
}

int rbrp_test_resident(int which) {
  return which == 0 ? rbrp_gemm_tn64_resident() : rbrp_gemm_tn32_resident();
}

int rbrp_test_gemm_nn64(void* stream, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double* C, int64_t ldc, int mode) {
  if (M < 0 || M % 4) return -2;
  if (N < 0) return -3;
  if (K < 0) return -4;
  if (lda < std::max<int64_t>(M, 1) || lda % 2) return -6;
  if (mode < GEMM_SET || mode > GEMM_ADD) return -11;
  if (M == 0 || N == 0 || K == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  // row ranges sized to the kernel's residency on this device, as the ID loop does
  HIPCHK(rbrp_launch_gemm_nn64(st, dim_static(M), M, dim_static(N), N, dim_static(K),
                               mat((double*)A, lda), mat((double*)B, ldb), mat(C, ldc), mode,
                               rbrp_gemm_nn64_resident()));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int rbrp_test_gemm_tn32(void* stream, int64_t M, int64_t N, int64_t K, const float* A,
                        const float* B, int64_t ld, double* C, int64_t ldc, int nsplit) {
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (K < 0 || K % 32) return -4;
  if (ld < std::max<int64_t>(K, 1) || ld % 4) return -7;
  if (nsplit < 1) return -10;
  if (M == 0 || N == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int64_t kc = (K + nsplit - 1) / nsplit;
  kc = std::max<int64_t>(32, (kc + 31) / 32 * 32);
  const int ns = (int)std::max<int64_t>(1, (K + kc - 1) / kc);
  double* P = nullptr;
  HIPCHK(hipMalloc(&P, sizeof(double) * M * N * ns));
  hipError_t e = rbrp_launch_gemm_tn32(st, dim_static(M), dim_static(N), M, N, K, ns, kc, A, 0,
                                       nullptr, B, 0, nullptr, ld, P, M, M * N);
  if (e == hipSuccess)
    e = rbrp_launch_reduce_splits(st, dim_static(M), dim_static(N), M, N, P, M, M * N, ns,
                                  mat(C, ldc), GEMM_SET);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(P);
  return (int)e;
}

int rbrp_test_gemm_upd32(void* stream, int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                         const float* Bt, int64_t ldbt, float* C, int64_t ldc, double* normpart) {
  if (M < 0 || (M & 1)) return -2;
  if (N < 0) return -3;
  if (K < 0 || K > 128) return -4;
  if ((lda & 1) || lda < std::max<int64_t>(M, 1)) return -6;
  if ((ldc & 1) || ldc < std::max<int64_t>(M, 1)) return -10;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(rbrp_launch_gemm_trail32(st, M, dim_static(N), N, dim_static(K), K, A, lda, 0, nullptr, Bt,
                                  ldbt, C, ldc, 0, nullptr, normpart, N));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int rbrp_test_panel(void* stream, const double* G, int64_t s, int64_t b, int filter, int64_t i,
                    const int32_t* J, int32_t* Jk, double* Rb, int32_t* Jin, int32_t* Jout,
                    int32_t* bt, int32_t* nswap, int32_t* status) {
  return rbrp_test_panel_inv(stream, G, s, b, filter, i, J, Jk, Rb, nullptr, Jin, Jout, bt, nswap,
                             status);
}

int rbrp_test_panel_inv(void* stream, const double* G, int64_t s, int64_t b, int filter,
                        int64_t i, const int32_t* J, int32_t* Jk, double* Rb, double* Rinv_out,
                        int32_t* Jin, int32_t* Jout, int32_t* bt, int32_t* nswap,
                        int32_t* status) {
  return rbrp_test_panel_kernel(stream, G, s, b, filter, i, J, Jk, Rb, Rinv_out, Jin, Jout, bt,
                                nswap, status, 0);
}

int rbrp_test_panel_kernel(void* stream, const double* G, int64_t s, int64_t b, int filter,
                           int64_t i, const int32_t* J, int32_t* Jk, double* Rb, double* Rinv_out,
                           int32_t* Jin, int32_t* Jout, int32_t* bt, int32_t* nswap,
                           int32_t* status, int kernel) {
  if (s < 1 || s > b) return -3;
  if (b < 1 || b > 128 || (kernel == 1 && b > 64)) return -4;
  if (kernel < 0 || kernel > 2) return -5;
  hipStream_t st = (hipStream_t)stream;
  RbrpBlockState* d = nullptr;
  double* Rinv = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(RbrpBlockState)));
  hipError_t e = hipMalloc(&Rinv, sizeof(double) * b * b);
  RbrpBlockState hs{};
  hs.s = (int)s;
  if (e == hipSuccess) e = hipMemcpyAsync(d, &hs, sizeof(hs), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = kernel == 1   ? rbrp_launch_pivchol_wave(st, G, s, (int)b, filter, i, J, d, Rb, Rinv, Jk,
                                                 Jin, Jout, nullptr)
        : kernel == 2 ? rbrp_launch_pivchol_regs(st, G, s, (int)b, filter, i, J, d, Rb, Rinv, Jk,
                                                 Jin, Jout, nullptr)
                      : rbrp_launch_pivchol_filter(st, G, s, (int)b, filter, i, J, d, Rb, Rinv, Jk,
                                                   Jin, Jout);
  if (e == hipSuccess) e = hipMemcpyAsync(&hs, d, sizeof(hs), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && Rinv_out)
    e = hipMemcpyAsync(Rinv_out, Rinv, sizeof(double) * b * b, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  (void)hipFree(Rinv);
  if (e != hipSuccess) return (int)e;
  *bt = hs.bt;
  *nswap = hs.nswap;
  *status = hs.status;
  return 0;
}

int rbrp_test_chol2(void* stream, const double* G2, int64_t bt, int64_t b, const double* Rb,
                    double* R2, double* R2inv, double* Rout, int32_t* status) {
  return rbrp_test_chol2_kernel(stream, G2, bt, b, Rb, R2, R2inv, Rout, status, 0);
}

int rbrp_test_chol2_kernel(void* stream, const double* G2, int64_t bt, int64_t b, const double* Rb,
                           double* R2, double* R2inv, double* Rout, int32_t* status, int kernel) {
  if (bt < 1 || bt > b) return -3;
  if (b < 1 || b > 128 || (kernel == 1 && b > 64)) return -4;
  if (kernel < 0 || kernel > 2) return -5;
  hipStream_t st = (hipStream_t)stream;
  RbrpBlockState* d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(RbrpBlockState)));
  RbrpBlockState hs{};
  hs.s = (int)bt;
  hs.bt = (int)bt;
  hipError_t e = hipMemcpyAsync(d, &hs, sizeof(hs), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = kernel == 1   ? rbrp_launch_chol2_wave(st, G2, bt, (int)b, d, Rb, R2, R2inv, Rout, bt)
        : kernel == 2 ? rbrp_launch_chol2_regs(st, G2, bt, (int)b, d, Rb, R2, R2inv, Rout, bt)
                      : rbrp_launch_chol2(st, G2, bt, (int)b, d, Rb, R2, R2inv, Rout, bt);
  if (e == hipSuccess) e = hipMemcpyAsync(&hs, d, sizeof(hs), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (e != hipSuccess) return (int)e;
  *status = hs.status;
  return 0;
}

int rbrp_test_sample(void* stream, const double* diags, int64_t i, int64_t n, int64_t kmax,
                     int64_t b, uint64_t seed, uint64_t offset, uint64_t draws_in, int32_t* J,
                     int32_t* s_out, uint64_t* draws_out) {
  if (n < 1 || i < 0 || i >= n) return -3;
  if (b < 1 || b > 256) return -6;
  hipStream_t st = (hipStream_t)stream;
  RbrpBlockState* d = nullptr;
  double* scratch = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(RbrpBlockState)));
  hipError_t e = hipMalloc(&scratch, sizeof(double) * 2 * n);
  RbrpBlockState hs{};
  hs.draws = draws_in;
  if (e == hipSuccess) e = hipMemcpyAsync(d, &hs, sizeof(hs), hipMemcpyHostToDevice, st);
  const uint64_t ctr[4] = {offset, 0, 0, 0};
  if (e == hipSuccess)
    e = rbrp_launch_sample(st, diags, i, n, kmax, (int)b, seed, 0, ctr, scratch, scratch + n, d, J);
  if (e == hipSuccess) e = hipMemcpyAsync(&hs, d, sizeof(hs), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  (void)hipFree(scratch);
  if (e != hipSuccess) return (int)e;
  *s_out = hs.s;
  *draws_out = hs.draws;
  return hs.status ? 100000 : 0;
}

int rbrp_test_sample_time(void* stream, const double* diags, int64_t n, int64_t b, uint64_t seed,
                          int reps, double* ms_per_call) {
  if (n < 1) return -3;
  if (b < 1 || b > 256) return -4;
  if (reps < 1) return -6;
  if (!ms_per_call) return -7;
  hipStream_t st = (hipStream_t)stream;
  RbrpBlockState* d = nullptr;
  double* scratch = nullptr;
  int* J = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(RbrpBlockState)));
  hipError_t e = hipMalloc(&scratch, sizeof(double) * 2 * n);
  if (e == hipSuccess) e = hipMalloc(&J, sizeof(int) * b);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  RbrpBlockState hs{};
  if (e == hipSuccess) e = hipMemcpyAsync(d, &hs, sizeof(hs), hipMemcpyHostToDevice, st);
  const uint64_t ctr[4] = {0, 0, 0, 0};
  // one untimed call, then reps calls between two events (the draws accumulate)
  if (e == hipSuccess) e = rbrp_launch_sample(st, diags, 0, n, n, (int)b, seed, 0, ctr, scratch, scratch + n, d, J);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    e = rbrp_launch_sample(st, diags, 0, n, n, (int)b, seed, 0, ctr, scratch, scratch + n, d, J);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) e = hipMemcpy(&hs, d, sizeof(hs), hipMemcpyDeviceToHost);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(d);
  (void)hipFree(scratch);
  (void)hipFree(J);
  if (e != hipSuccess) return (int)e;
  *ms_per_call = ms / reps;
  return hs.status ? 100000 : 0;
}

}  // extern "C"
