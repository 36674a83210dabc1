// rbrp.hpp -- C++ call surface of the reference (gpu/rbrp.hpp, gpu/rid.hpp) over the C ABI.
//
// Header-only: every function is an inline wrapper around include/rbrp.h, so a
// caller of the reference library recompiles against this header and links
// librbrp.so.  Errors are reported as return codes of the C ABI; the wrappers
// keep the reference's void signatures and return the status separately
// where the reference had none (RBRP_last_status()).
#ifndef RBRP_HPP_
#define RBRP_HPP_

#include <hip/hip_runtime.h>

#include <vector>

#include "rbrp.h"

namespace rbrp_detail {
inline int& last_status() {
  static thread_local int s = 0;
  return s;
}
// Householder QR flop count of an m x n matrix (LAPACK Working Note 41
// operation counts, as tabulated in gpu/flops.hpp:31-40).
inline double flops_dgeqrf(double m, double n) {
  const double muls = (m > n) ? n * (n * (0.5 - n / 3.0 + m) + m + 23.0 / 6.0)
                              : m * (m * (-0.5 - m / 3.0 + n) + 2.0 * n + 23.0 / 6.0);
  const double adds = (m > n) ? n * (n * (0.5 - n / 3.0 + m) + 5.0 / 6.0)
                              : m * (m * (-0.5 - m / 3.0 + n) + n + 5.0 / 6.0);
  return muls + adds;
}
}  // namespace rbrp_detail

// Status of the last wrapper call on this thread (0 = success; see rbrp.h).
inline int RBRP_last_status() { return rbrp_detail::last_status(); }

// gpu/rbrp.hpp:6-8.  A is an m x n column-major FP64 device pointer (lda = m),
// left unchanged.  Runs to min(m, n) with no tolerance, like the reference.
// ranks / error / t_ranks are cumulative per block; flops is the algorithmic
// FP64 count of the trailing updates (the reference declares it but never sets it).
inline void RBRP(int m, int n, double* A, int blk, std::vector<int>& ranks,
                 std::vector<double>& error, std::vector<double>& t_ranks, double& t_rbrp,
                 double& flops) {
  ranks.clear();
  error.clear();
  t_ranks.clear();
  t_rbrp = 0.0;
  flops = 0.0;
  rbrp_handle_t h = nullptr;
  int st = rbrp_create(&h, nullptr);
  if (st) {
    rbrp_detail::last_status() = st;
    return;
  }
  int32_t* perm = nullptr;
  if (hipMalloc(&perm, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)) != hipSuccess) {
    rbrp_destroy(h);
    rbrp_detail::last_status() = RBRP_ERR_ALLOC;
    return;
  }
  const int p = m < n ? m : n;
  std::vector<int64_t> r64((size_t)p + 1);
  error.resize((size_t)p + 1);
  t_ranks.resize((size_t)p + 1);
  rbrp_opts_t o;
  rbrp_opts_default(&o);
  o.b = blk;
  rbrp_result_t res{};
  res.perm = perm;
  res.X = nullptr;
  res.trace_cap = p + 1;
  res.ranks = r64.data();
  res.error = error.data();
  res.t_ranks = t_ranks.data();
  st = rbrp_colid_d(h, A, m, n, m, &o, &res);
  (void)hipFree(perm);
  rbrp_destroy(h);
  rbrp_detail::last_status() = st;
  const size_t nb = st ? 0 : (size_t)res.nblocks;
  error.resize(nb);
  t_ranks.resize(nb);
  ranks.resize(nb);
  int64_t prev = 0;
  for (size_t t = 0; t < nb; ++t) {
    ranks[t] = (int)r64[t];
    flops += 4.0 * (double)m * (double)(r64[t] - prev) * (double)(n - r64[t]);
    prev = r64[t];
  }
  t_rbrp = st ? 0.0 : res.t_total;
}

// Added overload with the tolerance and the skeleton / redundant sets and X.
// sk (k entries) and rd (n - k entries) are host vectors; X is a device buffer
// of k x (n - k) doubles (leading dimension k) allocated with hipMalloc that
// the caller releases with hipFree.  Returns the C-ABI status.
inline int RBRP(int m, int n, const double* A, int blk, double tol, std::vector<int>& sk,
                std::vector<int>& rd, double*& X, std::vector<int>& ranks,
                std::vector<double>& error, uint64_t seed = 0) {
  rbrp_handle_t h = nullptr;
  int st = rbrp_create(&h, nullptr);
  if (st) return st;
  const int p = m < n ? m : n;
  int32_t* perm = nullptr;
  double* xbuf = nullptr;
  st = (int)hipMalloc(&perm, sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!st) st = (int)hipMalloc(&xbuf, sizeof(double) * (size_t)(p > 0 ? p : 1) * (size_t)(n > 0 ? n : 1));
  std::vector<int64_t> r64((size_t)p + 1);
  std::vector<double> err((size_t)p + 1), tr((size_t)p + 1);
  rbrp_result_t res{};
  if (!st) {
    rbrp_opts_t o;
    rbrp_opts_default(&o);
    o.b = blk;
    o.tol = tol;
    o.seed = seed;
    res.perm = perm;
    res.X = xbuf;
    res.ldx = p > 0 ? p : 1;
    res.trace_cap = p + 1;
    res.ranks = r64.data();
    res.error = err.data();
    res.t_ranks = tr.data();
    st = rbrp_colid_d(h, A, m, n, m, &o, &res);
  }
  if (!st) {
    const int k = (int)res.k;
    std::vector<int32_t> hp((size_t)n);
    st = (int)hipMemcpy(hp.data(), perm, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost);
    sk.assign(hp.begin(), hp.begin() + k);
    rd.assign(hp.begin() + k, hp.end());
    X = nullptr;
    if (!st && k > 0 && n > k) {
      st = (int)hipMalloc(&X, sizeof(double) * (size_t)k * (size_t)(n - k));
      if (!st)
        st = (int)hipMemcpy2D(X, sizeof(double) * (size_t)k, xbuf, sizeof(double) * (size_t)res.ldx,
                              sizeof(double) * (size_t)k, (size_t)(n - k), hipMemcpyDeviceToDevice);
    }
    ranks.assign(r64.begin(), r64.begin() + res.nblocks);
    error.assign(err.begin(), err.begin() + res.nblocks);
  }
  if (perm) (void)hipFree(perm);
  if (xbuf) (void)hipFree(xbuf);
  rbrp_destroy(h);
  rbrp_detail::last_status() = st;
  return st;
}

// gpu/rid.hpp:19-20.  h_sk (rank) and h_rd (n - rank) are host arrays the
// caller provides, T (rank x (n - rank), leading dimension rank) is allocated
// here with hipMalloc and released by the caller with hipFree.  flops is the
// reference's count 2 m n (2r) + FLOPS_DGEQRF(2r, n) + (2r)^2 (n - 2r)
// (randCPQR_OS.cu:199).  The sketch is drawn from the Philox stream (seed, 0);
// A is m x n with leading dimension m.
inline void RandCPQR_OS(const double* A, int m, int n, int rank, int* h_sk, int* h_rd, double*& T,
                        double& flops, uint64_t seed = 0) {
  T = nullptr;
  flops = 0.0;
  rbrp_handle_t h = nullptr;
  int st = rbrp_create(&h, nullptr);
  if (!st && n > rank) st = (int)hipMalloc(&T, sizeof(double) * (size_t)rank * (size_t)(n - rank));
  if (!st)
    st = rbrp_randcpqr_os_d(h, A, m, n, m, rank, seed, 0, nullptr, 0, h_sk, h_rd, T, rank);
  if (h) rbrp_destroy(h);
  rbrp_detail::last_status() = st;
  const double k = 2.0 * rank, mm = m, nn = n;
  flops = 2.0 * mm * nn * k + rbrp_detail::flops_dgeqrf(k, nn) + k * k * (nn - k);
}

// gpu/id_cpqr.hpp:6-8.  A is an m x n column-major FP64 device pointer (lda = m),
// left unchanged.  Full column-pivoted QR of A, then for each rank r in ranks
// (1 <= r <= min(m, n); the reference asserts 0 < r < min(m, n)): err = the
// squared relative error of the CPQR ID of rank r, t_ranks = the reference's
// partial-factorization time model plus the solve time, t_cpqr = the
// factorization's seconds, flops = FLOPS_DGEQRF(m, n) + sum r^2 (n - r).
inline void CPQR(int m, int n, double* A, const std::vector<int> ranks, std::vector<double>& err,
                 std::vector<double>& t_ranks, double& t_cpqr, double& flops) {
  err.assign(ranks.size(), 0.0);
  t_ranks.assign(ranks.size(), 0.0);
  t_cpqr = 0.0;
  flops = 0.0;
  rbrp_handle_t h = nullptr;
  int st = rbrp_create(&h, nullptr);
  if (!st)
    st = rbrp_cpqr_id_d(h, A, m, n, m, (int64_t)ranks.size(), ranks.data(), err.data(),
                        t_ranks.data(), &t_cpqr, &flops, nullptr);
  if (h) rbrp_destroy(h);
  rbrp_detail::last_status() = st;
}

// gpu/rid.hpp:16-17.  As RandCPQR_OS without oversampling: the sketch G is
// rank x m and T comes from its own pivoted-QR factor.  flops as
// randCPQR.cu:67: 2 m n r + FLOPS_DGEQRF(r, n) + r^2 (n - r).
inline void RandCPQR(const double* A, int m, int n, int rank, int* h_sk, int* h_rd, double*& T,
                     double& flops, uint64_t seed = 0) {
  T = nullptr;
  flops = 0.0;
  rbrp_handle_t h = nullptr;
  int st = rbrp_create(&h, nullptr);
  if (!st && n > rank) st = (int)hipMalloc(&T, sizeof(double) * (size_t)rank * (size_t)(n - rank));
  if (!st) st = rbrp_randcpqr_d(h, A, m, n, m, rank, seed, 0, nullptr, 0, h_sk, h_rd, T, rank);
  if (h) rbrp_destroy(h);
  rbrp_detail::last_status() = st;
  const double k = rank, mm = m, nn = n;
  flops = 2.0 * mm * nn * k + rbrp_detail::flops_dgeqrf(k, nn) + k * k * (nn - k);
}

// gpu/rid.hpp:9-10.  As RandLUPP_OS without oversampling (G is rank x m,
// T = (L21 L11^{-1})^T); sk, rd and T are allocated here with hipMalloc and
// released by the caller with hipFree.  flops as randLUPP.cu:109.
inline void RandLUPP(const double* A, int m, int n, int rank, int*& sk, int*& rd, double*& T,
                     double& flops, uint64_t seed = 0) {
  sk = rd = nullptr;
  T = nullptr;
  flops = 0.0;
  rbrp_handle_t h = nullptr;
  int st = rbrp_create(&h, nullptr);
  if (!st) st = (int)hipMalloc(&sk, sizeof(int) * (size_t)rank);
  if (!st && n > rank) st = (int)hipMalloc(&rd, sizeof(int) * (size_t)(n - rank));
  if (!st && n > rank) st = (int)hipMalloc(&T, sizeof(double) * (size_t)rank * (size_t)(n - rank));
  if (!st)
    st = rbrp_randlupp_d(h, A, m, n, m, rank, seed, 0, nullptr, 0, (int32_t*)sk, (int32_t*)rd, T, rank);
  if (h) rbrp_destroy(h);
  rbrp_detail::last_status() = st;
  const double k = rank, mm = m, nn = n;
  flops = 2.0 * mm * nn * k + 2.0 * mm * k * k / 3.0 + k * k * (mm - k);
}

// gpu/rid.hpp:12-13.  sk, rd (device int arrays) and T (device, rank x
// (n - rank), leading dimension rank) are allocated here with hipMalloc; the
// caller frees them with hipFree.  flops as randLUPP_OS.cu:100.
inline void RandLUPP_OS(const double* A, int m, int n, int rank, int*& sk, int*& rd, double*& T,
                        double& flops, uint64_t seed = 0) {
  sk = rd = nullptr;
  T = nullptr;
  flops = 0.0;
  rbrp_handle_t h = nullptr;
  int st = rbrp_create(&h, nullptr);
  if (!st) st = (int)hipMalloc(&sk, sizeof(int) * (size_t)rank);
  if (!st && n > rank) st = (int)hipMalloc(&rd, sizeof(int) * (size_t)(n - rank));
  if (!st && n > rank) st = (int)hipMalloc(&T, sizeof(double) * (size_t)rank * (size_t)(n - rank));
  if (!st)
    st = rbrp_randlupp_os_d(h, A, m, n, m, rank, seed, 0, nullptr, 0, (int32_t*)sk, (int32_t*)rd, T,
                            rank);
  if (h) rbrp_destroy(h);
  rbrp_detail::last_status() = st;
  const double k = rank, mm = m, nn = n;
  flops = 2.0 * mm * nn * k + 2.0 * mm * k * k / 3.0 + k * k * (mm - k);
}

// gpu/rid.hpp:5-7 (declared there, defined nowhere under gpu/; the semantics
// are blocked_rlupp_id of python/cache.ipynb c8:80-143).  Adaptive rank: sk,
// rd (device int arrays of rank and n - rank entries) and T (device, rank x
// (n - rank), leading dimension rank) are allocated here with hipMalloc and
// released by the caller with hipFree; tol is relative to ||A||_F.
inline void RandAdapLUPP(const double* A, int m, int n, int*& sk, int*& rd, double*& T, int& rank,
                         double& flops, double tol = 1e-8, int blk = 16, uint64_t seed = 0) {
  sk = rd = nullptr;
  T = nullptr;
  rank = 0;
  flops = 0.0;
  rbrp_handle_t h = nullptr;
  int32_t *wsk = nullptr, *wrd = nullptr;
  double* wT = nullptr;
  const size_t p = (size_t)(m < n ? m : n);
  int st = rbrp_create(&h, nullptr);
  if (!st) st = (int)hipMalloc(&wsk, sizeof(int32_t) * (size_t)n);
  if (!st) st = (int)hipMalloc(&wrd, sizeof(int32_t) * (size_t)n);
  if (!st) st = (int)hipMalloc(&wT, sizeof(double) * p * (size_t)n);
  int64_t k = 0;
  if (!st)
    st = rbrp_randadaplupp_d(h, A, m, n, m, tol, blk, seed, 0, wsk, wrd, wT, (int64_t)p, &k, nullptr,
                             &flops);
  if (!st) {
    // exact-size outputs (T repacked with leading dimension rank)
    rank = (int)k;
    st = (int)hipMalloc(&sk, sizeof(int) * (size_t)(k > 0 ? k : 1));
    if (!st) st = (int)hipMemcpy(sk, wsk, sizeof(int) * (size_t)k, hipMemcpyDeviceToDevice);
    if (!st && n > k) st = (int)hipMalloc(&rd, sizeof(int) * (size_t)(n - k));
    if (!st && n > k)
      st = (int)hipMemcpy(rd, wrd, sizeof(int) * (size_t)(n - k), hipMemcpyDeviceToDevice);
    if (!st && n > k && k > 0) st = (int)hipMalloc(&T, sizeof(double) * (size_t)k * (size_t)(n - k));
    if (!st && n > k && k > 0)
      st = (int)hipMemcpy2D(T, sizeof(double) * (size_t)k, wT, sizeof(double) * p, sizeof(double) * (size_t)k,
                            (size_t)(n - k), hipMemcpyDeviceToDevice);
  }
  if (wsk) (void)hipFree(wsk);
  if (wrd) (void)hipFree(wrd);
  if (wT) (void)hipFree(wT);
  if (h) rbrp_destroy(h);
  rbrp_detail::last_status() = st;
}

#endif  // RBRP_HPP_
