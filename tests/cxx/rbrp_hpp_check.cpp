// Drop-in check of include/rbrp.hpp: calls the reference's RBRP signature and
// the added tolerance overload on a device matrix and prints one JSON line.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "rbrp.hpp"

int main() {
  const int m = 400, n = 300;
  // A = sum_k exp(-k/8) u_k v_k^T with random u, v (rapidly decaying rank)
  std::mt19937_64 g(7);
  std::normal_distribution<double> nd;
  std::vector<double> A((size_t)m * n, 0.0);
  for (int k = 0; k < 80; ++k) {
    std::vector<double> u(m), v(n);
    for (auto& x : u) x = nd(g);
    for (auto& x : v) x = nd(g);
    const double s = std::exp(-k / 8.0);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) A[i + (size_t)j * m] += s * u[i] * v[j];
  }
  double* dA = nullptr;
  if (hipMalloc(&dA, sizeof(double) * A.size()) != hipSuccess) return 2;
  if (hipMemcpy(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice) != hipSuccess) return 2;
  std::vector<int> ranks;
  std::vector<double> error, t_ranks;
  double t_rbrp = 0, flops = 0;
  RBRP(m, n, dA, 32, ranks, error, t_ranks, t_rbrp, flops);
  const int st1 = RBRP_last_status();
  std::vector<int> sk, rd, r2;
  std::vector<double> e2;
  double* X = nullptr;
  const int st2 = RBRP(m, n, dA, 32, 1e-6, sk, rd, X, r2, e2, 3);
  // host check of the ID error ||A(:,rd) - A(:,sk) X|| / ||A||
  const int k = (int)sk.size();
  std::vector<double> hX((size_t)k * (n - k));
  if (X) (void)hipMemcpy(hX.data(), X, sizeof(double) * hX.size(), hipMemcpyDeviceToHost);
  double num = 0, den = 0;
  for (int j = 0; j < n - k; ++j)
    for (int i = 0; i < m; ++i) {
      double v = A[i + (size_t)rd[j] * m];
      for (int q = 0; q < k; ++q) v -= A[i + (size_t)sk[q] * m] * hX[q + (size_t)j * k];
      num += v * v;
    }
  for (double x : A) den += x * x;
  // sketch comparators at the rank the ID found
  std::vector<int> csk((size_t)k), crd((size_t)(n - k));
  double* cT = nullptr;
  double cfl = 0;
  RandCPQR_OS(dA, m, n, k, csk.data(), crd.data(), cT, cfl, 5);
  const int st3 = RBRP_last_status();
  int *lsk = nullptr, *lrd = nullptr;
  double* lT = nullptr;
  double lfl = 0;
  RandLUPP_OS(dA, m, n, k, lsk, lrd, lT, lfl, 6);
  const int st4 = RBRP_last_status();
  auto id_err = [&](const std::vector<int>& s_, const std::vector<int>& r_, const double* dT) {
    std::vector<double> hT((size_t)k * (n - k));
    (void)hipMemcpy(hT.data(), dT, sizeof(double) * hT.size(), hipMemcpyDeviceToHost);
    double num2 = 0;
    for (int j = 0; j < n - k; ++j)
      for (int i = 0; i < m; ++i) {
        double v = A[i + (size_t)r_[j] * m];
        for (int q = 0; q < k; ++q) v -= A[i + (size_t)s_[q] * m] * hT[q + (size_t)j * k];
        num2 += v * v;
      }
    return std::sqrt(num2 / den);
  };
  std::vector<int> lskh((size_t)k), lrdh((size_t)(n - k));
  (void)hipMemcpy(lskh.data(), lsk, sizeof(int) * k, hipMemcpyDeviceToHost);
  (void)hipMemcpy(lrdh.data(), lrd, sizeof(int) * (n - k), hipMemcpyDeviceToHost);
  const double ecp = (st3 == 0) ? id_err(csk, crd, cT) : -1.0;
  const double elu = (st4 == 0) ? id_err(lskh, lrdh, lT) : -1.0;
  if (cT) (void)hipFree(cT);
  if (lT) (void)hipFree(lT);
  if (lsk) (void)hipFree(lsk);
  if (lrd) (void)hipFree(lrd);
  // adaptive randomised LUPP ID (gpu/rid.hpp:5-7) at a relative tolerance
  int *ask = nullptr, *ard = nullptr, arank = 0;
  double* aT = nullptr;
  double afl = 0;
  RandAdapLUPP(dA, m, n, ask, ard, aT, arank, afl, 1e-4, 16, 7);
  const int st6 = RBRP_last_status();
  double eada = -1.0;
  if (st6 == 0 && arank > 0 && arank < n) {
    std::vector<int> ah((size_t)arank), arh((size_t)(n - arank));
    (void)hipMemcpy(ah.data(), ask, sizeof(int) * arank, hipMemcpyDeviceToHost);
    (void)hipMemcpy(arh.data(), ard, sizeof(int) * (n - arank), hipMemcpyDeviceToHost);
    std::vector<double> hT((size_t)arank * (n - arank));
    (void)hipMemcpy(hT.data(), aT, sizeof(double) * hT.size(), hipMemcpyDeviceToHost);
    double num2 = 0;
    for (int j = 0; j < n - arank; ++j)
      for (int i = 0; i < m; ++i) {
        double v = A[i + (size_t)arh[j] * m];
        for (int q = 0; q < arank; ++q) v -= A[i + (size_t)ah[q] * m] * hT[q + (size_t)j * arank];
        num2 += v * v;
      }
    eada = std::sqrt(num2 / den);
  }
  if (ask) (void)hipFree(ask);
  if (ard) (void)hipFree(ard);
  if (aT) (void)hipFree(aT);
  std::printf("{\"status6\": %d, \"adap_rank\": %d, \"adap_error\": %.6e, \"adap_flops\": %.6e}\n",
              st6, arank, eada, afl);
  // full-matrix CPQR ID at half the rank and at the rank the ID found
  std::vector<double> qerr, qt;
  double tq = 0, qfl = 0;
  CPQR(m, n, dA, std::vector<int>{k / 2 > 0 ? k / 2 : 1, k}, qerr, qt, tq, qfl);
  const int st5 = RBRP_last_status();
  std::printf("{\"status5\": %d, \"cpqr_sq_errors\": [%.6e, %.6e], \"t_cpqr\": %.6e, "
              "\"t_ranks\": [%.6e, %.6e], \"cpqr_flops\": %.6e}\n",
              st5, qerr.size() > 1 ? qerr[0] : -1.0, qerr.size() > 1 ? qerr[1] : -1.0, tq,
              qt.size() > 1 ? qt[0] : -1.0, qt.size() > 1 ? qt[1] : -1.0, qfl);
  std::printf("{\"status3\": %d, \"status4\": %d, \"cpqr_os_error\": %.6e, \"lupp_os_error\": %.6e}\n",
              st3, st4, ecp, elu);
  std::printf("{\"status1\": %d, \"final_rank\": %d, \"nblocks\": %zu, \"status2\": %d, \"k\": %d, "
              "\"id_error\": %.6e, \"flops\": %.6e}\n",
              st1, ranks.empty() ? 0 : ranks.back(), ranks.size(), st2, k, std::sqrt(num / den), flops);
  if (X) (void)hipFree(X);
  (void)hipFree(dA);
  return (st1 == 0 && st2 == 0 && st3 == 0 && st4 == 0 && st5 == 0 && st6 == 0) ? 0 : 1;
}
