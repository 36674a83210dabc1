// Full-matrix column-pivoted QR ID baseline behind the C ABI: the "CPQR" line
// of the reference's experiments, CPQR of gpu/id_cpqr.hpp:6-8 and
// gpu/id_cpqr.cu:30-172.
//
// A working copy of A is factored by the blocked Householder QR with column
// pivoting (rbrp_qrcp: dgeqp3's pivot rule on downdated partial column
// norms, 32-column panels with the trailing update on MFMA) to min(m, n)
// steps.  For each
// requested rank r: T = R11^{-1} R12 by the blocked TRSM, and the squared
// relative error ||A(:, P(r:n)) - A(:, P(0:r)) T||_F^2 / ||A||_F^2, with the
// permuted columns read through the pivot list rather than a permuted copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/rbrp.h"
#include "rbrp_handle.h"
#include "rbrp_kernels.h"

#define HIPCHK(x)                         \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) return (int)e_; \
  } while (0)

namespace {

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Work of the first r steps of a pivoted QR of an m x n matrix, counting step
// j (0-based) as (m - j)(n - j) element updates: with q = min(m, n) and
// p = max(m, n) - q, the step with t = q - j columns left costs (p + t) t.
// The model of gpu/id_cpqr.cu:93-104 charges a partial factorization to rank
// r the fraction work(r) / work(q) of the full one.
double qr_work_from(int64_t m, int64_t n, int64_t tmin) {
  // sum over t = tmin .. q of (p + t) t
  const double q = (double)std::min(m, n), p = (double)std::max(m, n) - q;
  auto upto = [&](double t) { return p * t * (t + 1) / 2.0 + t * (t + 1) * (2 * t + 1) / 6.0; };
  return upto(q) - upto((double)tmin - 1.0);
}

// Householder QR flop count of an m x n matrix: the LAPACK Working Note 41
// operation counts that gpu/flops.hpp:31-40 tabulates (multiplies + adds).
double flops_dgeqrf(double m, double n) {
  const double muls = (m > n) ? n * (n * (0.5 - n / 3.0 + m) + m + 23.0 / 6.0)
                              : m * (m * (-0.5 - m / 3.0 + n) + 2.0 * n + 23.0 / 6.0);
  const double adds = (m > n) ? n * (n * (0.5 - n / 3.0 + m) + 5.0 / 6.0)
                              : m * (m * (-0.5 - m / 3.0 + n) + n + 5.0 / 6.0);
  return muls + adds;
}

// Sum of a small host vector in index order.
double host_sum(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x;
  return s;
}

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  hipError_t create() {
    hipError_t e = hipEventCreate(&a);
    return e != hipSuccess ? e : hipEventCreate(&b);
  }
  // seconds between the two records, after waiting for the second
  hipError_t seconds(double* s) const {
    hipError_t e = hipEventSynchronize(b);
    if (e != hipSuccess) return e;
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, a, b);
    *s = 1e-3 * ms;
    return e;
  }
};

}  // namespace

extern "C" int rbrp_cpqr_id_d(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda,
                              int64_t nranks, const int32_t* ranks, double* error, double* t_ranks,
                              double* t_cpqr, double* flops, int32_t* d_perm) {
  if (!h) return -1;
  if (m < 1) return -3;
  if (n < 1) return -4;
  if (lda < m) return -5;
  if (!dA) return -2;
  if (nranks < 0) return -6;
  if (nranks > 0 && !ranks) return -7;
  if (nranks > 0 && !error) return -8;
  if (nranks > 0 && !t_ranks) return -9;
  if (!t_cpqr) return -10;
  if (!flops) return -11;
  if (m > 0x7fffffffLL || n > 0x7fffffffLL) return -3;
  const int64_t q = std::min(m, n);
  for (int64_t t = 0; t < nranks; ++t)
    if (ranks[t] < 1 || ranks[t] > q) return -7;

  hipStream_t st = h->stream;
  const size_t bW = align256(sizeof(double) * m * n), bE = bW;
  const size_t bT = align256(sizeof(double) * q * n);
  const size_t bDinv = align256(sizeof(double) * 4 * ((q + 255) / 256) * 4096);
  const size_t bvn = align256(sizeof(double) * n), btau = align256(sizeof(double) * q);
  const size_t bjp = align256(sizeof(int) * n);
  const size_t bqp = align256(rbrp_qrcp_work_bytes(m, n));
  void* arena = nullptr;
  HIPCHK(rbrp_sketch_arena(h, bW + bE + bT + bDinv + bvn + btau + bjp + bqp, &arena));
  char* p = (char*)arena;
  double* W = (double*)p;
  p += bW;
  double* E = (double*)p;
  p += bE;
  double* T = (double*)p;
  p += bT;
  double* Dinv = (double*)p;
  p += bDinv;
  double* cn = (double*)p;
  p += bvn;
  double* tau = (double*)p;
  p += btau;
  int* jpvt = (int*)p;
  p += bjp;
  void* qwork = p;

  // ||A||_F^2 before the timed factorization (id_cpqr.cu:37-38)
  std::vector<double> hcn((size_t)n);
  HIPCHK(rbrp_launch_colnorm2(st, m, n, dA, lda, cn));
  HIPCHK(hipMemcpyAsync(hcn.data(), cn, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double anrm2 = host_sum(hcn);

  EventPair ev;
  HIPCHK(ev.create());
  HIPCHK(hipEventRecord(ev.a, st));
  HIPCHK(rbrp_launch_copy_block(st, m, n, dA, lda, W, m));
  HIPCHK(rbrp_qrcp(st, m, n, W, m, jpvt, tau, qwork, h->host_scalar));
  HIPCHK(hipEventRecord(ev.b, st));
  double tq = 0.0;
  HIPCHK(ev.seconds(&tq));
  *t_cpqr = tq;
  double fl = flops_dgeqrf((double)m, (double)n);
  if (d_perm) HIPCHK(hipMemcpyAsync(d_perm, jpvt, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));

  const double wq = qr_work_from(m, n, 1);
  std::vector<double> hen;
  for (int64_t t = 0; t < nranks; ++t) {
    const int64_t r = ranks[t], nc = n - r;
    // the first r steps of the factorization (id_cpqr.cu:93-104)
    const double tr = wq > 0.0 ? tq * qr_work_from(m, n, q - r + 1) / wq : 0.0;
    if (nc == 0) {
      error[t] = 0.0;
      t_ranks[t] = tr;
      continue;
    }
    // T = R11^{-1} R12 (id_cpqr.cu:118-136), timed as the reference does
    HIPCHK(rbrp_launch_copy_block(st, r, nc, W + r * m, m, T, r));
    HIPCHK(hipEventRecord(ev.a, st));
    HIPCHK(rbrp_trsm_upper_left(st, r, nc, W, m, T, r, nullptr, Dinv));
    HIPCHK(hipEventRecord(ev.b, st));
    double ts = 0.0;
    HIPCHK(ev.seconds(&ts));
    t_ranks[t] = tr + ts;
    fl += (double)r * r * nc;
    // E = A(:, P(r:n)) - A(:, P(0:r)) T and its squared norm (id_cpqr.cu:142-166)
    HIPCHK(rbrp_launch_gather_cols(st, m, nc, dA, lda, jpvt + r, E, m));
    HIPCHK(rbrp_launch_gemm_nn(st, dim_static(m), m, dim_static(nc), nc, dim_static(r),
                               mat_cols((double*)dA, lda, jpvt), mat(T, r), mat(E, m), GEMM_SUB,
                               nullptr, 0));
    HIPCHK(rbrp_launch_colnorm2(st, m, nc, E, m, cn));
    hen.resize((size_t)nc);
    HIPCHK(hipMemcpyAsync(hen.data(), cn, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    error[t] = anrm2 > 0.0 ? host_sum(hen) / anrm2 : 0.0;
  }
  *flops = fl;
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}
