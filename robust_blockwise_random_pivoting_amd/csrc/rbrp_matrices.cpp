// Test matrices of the reference's driver (gpu/driver.cpp:23-111), generated on
// the host (no GPU needed) from a seeded std::mt19937 stream; the reference
// draws from unseeded generators (Eigen's Random, std::random_device), so the
// families are the same and the individual matrices are not.
//
//   GE  (driver.cpp:23-47):  A = U diag(s) V^T, n x n, with U, V the Q factors of
//        uniform [-1, 1] matrices and s_i = 1 for i < 100,
//        max(0.8^(i - 100), 1e-5) after.
//   SNN (driver.cpp:50-91):  A = (U D) V, n x n, with U and V sparse (ceil(s n)
//        random entries per column on average, uniform [0, 1) values,
//        duplicates summed) and D = diag(10/100 for i < 100, 1/(i - 99) after).
//   GMM (driver.cpp:93-111): A (d x n) = N(0, 1) entries, plus 10 (j + 1) on
//        row j of the columns of cluster j (100 clusters of n / 100 columns).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/rbrp.h"

namespace {

// Explicit Q (n x n, column-major) of the Householder QR of M (n x n, overwritten).
void householder_q(int64_t n, std::vector<double>& M, std::vector<double>& Q) {
  std::vector<double> tau((size_t)n, 0.0);
  for (int64_t j = 0; j < n; ++j) {
    double* x = &M[(size_t)j * n];
    double xn2 = 0.0;
    for (int64_t i = j + 1; i < n; ++i) xn2 += x[i] * x[i];
    if (xn2 == 0.0) continue;
    const double alpha = x[j], nrm = std::sqrt(alpha * alpha + xn2);
    const double beta = alpha >= 0.0 ? -nrm : nrm;
    tau[(size_t)j] = (beta - alpha) / beta;
    const double sc = 1.0 / (alpha - beta);
    for (int64_t i = j + 1; i < n; ++i) x[i] *= sc;
    x[j] = beta;
    // apply H_j = I - tau v v^T (v_j = 1) to the columns right of j
    for (int64_t c = j + 1; c < n; ++c) {
      double* w = &M[(size_t)c * n];
      double dot = w[j];
      for (int64_t i = j + 1; i < n; ++i) dot += x[i] * w[i];
      const double f = tau[(size_t)j] * dot;
      w[j] -= f;
      for (int64_t i = j + 1; i < n; ++i) w[i] -= f * x[i];
    }
  }
  // Q = H_0 H_1 ... H_{n-1} I, applied from the last reflector back
  Q.assign((size_t)n * n, 0.0);
  for (int64_t i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  for (int64_t j = n - 1; j >= 0; --j) {
    const double t = tau[(size_t)j];
    if (t == 0.0) continue;
    const double* v = &M[(size_t)j * n];
    for (int64_t c = j; c < n; ++c) {
      double* w = &Q[(size_t)c * n];
      double dot = w[j];
      for (int64_t i = j + 1; i < n; ++i) dot += v[i] * w[i];
      const double f = t * dot;
      w[j] -= f;
      for (int64_t i = j + 1; i < n; ++i) w[i] -= f * v[i];
    }
  }
}

}  // namespace

extern "C" {

int rbrp_matrix_ge_h(int64_t n, uint64_t seed, double* A) {
  if (n < 1) return -1;
  if (!A) return -3;
  std::mt19937 rng((uint32_t)seed);
  std::uniform_real_distribution<double> unif(-1.0, 1.0);
  std::vector<double> M((size_t)n * n), U, V;
  for (double& x : M) x = unif(rng);
  householder_q(n, M, U);
  for (double& x : M) x = unif(rng);
  householder_q(n, M, V);
  std::vector<double> s((size_t)n);
  for (int64_t i = 0; i < n; ++i)
    s[(size_t)i] = i < 100 ? 1.0 : std::max(std::pow(0.8, (double)(i - 100)), 1e-5);
  // A = (U diag(s)) V^T: column c of A = sum_k U(:, k) s_k V(c, k)
  for (int64_t c = 0; c < n; ++c) {
    double* a = &A[(size_t)c * n];
    for (int64_t i = 0; i < n; ++i) a[i] = 0.0;
    for (int64_t k = 0; k < n; ++k) {
      const double f = s[(size_t)k] * V[(size_t)k * n + c];
      const double* u = &U[(size_t)k * n];
      for (int64_t i = 0; i < n; ++i) a[i] += f * u[i];
    }
  }
  return 0;
}

int rbrp_matrix_snn_h(int64_t n, double s, uint64_t seed, double* A) {
  if (n < 1) return -1;
  if (!(s > 0.0) || !(s <= 1.0)) return -2;
  if (!A) return -4;
  const int64_t m = 100;
  std::vector<double> d((size_t)n);
  for (int64_t i = 0; i < n; ++i) d[(size_t)i] = i < m ? 10.0 / m : 1.0 / (double)(i - m + 1);
  const int64_t nnz = (int64_t)std::ceil(s * (double)n);
  std::mt19937 rng((uint32_t)seed);
  std::uniform_int_distribution<int> uni(0, (int)n - 1);
  std::uniform_real_distribution<double> unf(0.0, 1.0);
  // U D kept column-wise as (row, value) lists; V as triplets (the draws
  // alternate between the two factors as driver.cpp:73-82 does)
  std::vector<std::vector<std::pair<int, double>>> ud((size_t)n);
  std::vector<int> vr, vc;
  std::vector<double> vv;
  vr.reserve((size_t)(n * nnz));
  vc.reserve((size_t)(n * nnz));
  vv.reserve((size_t)(n * nnz));
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = 0; k < nnz; ++k) {
      int r = uni(rng), c = uni(rng);
      const double u = unf(rng);
      ud[(size_t)c].emplace_back(r, u * d[(size_t)c]);
      r = uni(rng);
      c = uni(rng);
      const double v = unf(rng);
      vr.push_back(r);
      vc.push_back(c);
      vv.push_back(v);
    }
  for (size_t i = 0; i < (size_t)n * n; ++i) A[i] = 0.0;
  // A(:, c) += (U D)(:, r) V(r, c) for every entry of V (duplicates add up)
  for (size_t t = 0; t < vv.size(); ++t) {
    double* a = &A[(size_t)vc[t] * n];
    for (const auto& e : ud[(size_t)vr[t]]) a[e.first] += e.second * vv[t];
  }
  return 0;
}

int rbrp_matrix_gmm_h(int64_t n, int64_t d, uint64_t seed, double* A) {
  const int64_t k = 100;  // clusters
  if (n < k) return -1;   // n / 100 columns per cluster must be at least one
  if (d < k) return -2;
  const int64_t m = n / k;
  if ((n - 1) / m >= d) return -2;  // the last cluster's shift row must exist
  if (!A) return -4;
  std::mt19937 rng((uint32_t)seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (size_t i = 0; i < (size_t)d * n; ++i) A[i] = nd(rng);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t j = i / m;
    A[(size_t)i * d + j] += 10.0 * (double)(j + 1);
  }
  return 0;
}

}  // extern "C"
