// rbrp_driver: the experiment driver of the reference (gpu/driver.cpp:114-306)
// over this library's C++ surface (include/rbrp.hpp).
//
// Generates a test matrix (GE, SNN or GMM; gpu/driver.cpp:23-111), runs RBRP
// to full rank with block size -bs and prints the ranks / error / time of its
// first -nb blocks, then the full-matrix CPQR ID, RandCPQR-OS and RandLUPP-OS
// at the same ranks, each as "error" and "time" lists, in the reference's
// layout (gpu/print.hpp:4-10): a blank line, "name:", the values separated by
// spaces, a blank line.  Errors are squared relative Frobenius errors, as the
// reference prints them.
//
// Flags of the reference: -n <columns> (1000), -bs <block size> (32),
// -nb <blocks to report> (6).  Added: -matrix ge|snn|gmm (gmm, the
// reference's active choice at driver.cpp:146), -d <GMM dimension> (1000),
// -s <SNN density> (0.1), -seed <generator seed> (1986).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "rbrp.hpp"

namespace {

template <typename T>
void print_list(const std::vector<T>& v, const std::string& name) {
  std::cout << std::endl << name << ":" << std::endl;
  for (const T& x : v) std::cout << x << " ";
  std::cout << std::endl << std::endl;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ||A(:, rd) - A(:, sk) T||_F^2 / ||A||_F^2 on the host (driver.cpp:259-261)
double id_error_sq(const std::vector<double>& A, int m, int n, const std::vector<int>& sk,
                   const std::vector<int>& rd, const std::vector<double>& T, double anrm2) {
  const int k = (int)sk.size();
  double num = 0.0;
  std::vector<double> col((size_t)m);
  for (int j = 0; j < n - k; ++j) {
    for (int i = 0; i < m; ++i) col[(size_t)i] = A[(size_t)rd[(size_t)j] * m + i];
    for (int q = 0; q < k; ++q) {
      const double t = T[(size_t)j * k + q];
      const double* a = &A[(size_t)sk[(size_t)q] * m];
      for (int i = 0; i < m; ++i) col[(size_t)i] -= a[i] * t;
    }
    for (int i = 0; i < m; ++i) num += col[(size_t)i] * col[(size_t)i];
  }
  return anrm2 > 0.0 ? num / anrm2 : 0.0;
}

int fail(const char* what, int st) {
  std::fprintf(stderr, "rbrp_driver: %s failed with status %d\n", what, st);
  return 1;
}

}  // namespace

int main(int argc, char* argv[]) {
  int n = 1000, bs = 32, nb = 6, d = 1000;
  double s = 0.1;
  unsigned long long seed = 1986;
  std::string matrix = "gmm";
  for (int i = 1; i < argc; ++i) {
    const bool more = i + 1 < argc;
    if (!std::strcmp(argv[i], "-n") && more) n = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "-bs") && more) bs = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "-nb") && more) nb = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "-d") && more) d = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "-s") && more) s = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "-seed") && more) seed = std::strtoull(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "-matrix") && more) matrix = argv[++i];
    else {
      std::fprintf(stderr,
                   "usage: rbrp_driver [-n N] [-bs B] [-nb NB] [-matrix ge|snn|gmm] [-d D] "
                   "[-s S] [-seed SEED]\n");
      return 2;
    }
  }
  if (n < 1 || bs < 1 || nb < 1) return fail("argument check", -1);

  // ---- test matrix (driver.cpp:137-153) ------------------------------------
  auto t0 = std::chrono::steady_clock::now();
  int m = n, st = 0;
  std::vector<double> A;
  if (matrix == "ge") {
    A.resize((size_t)n * n);
    st = rbrp_matrix_ge_h(n, seed, A.data());
  } else if (matrix == "snn") {
    A.resize((size_t)n * n);
    st = rbrp_matrix_snn_h(n, s, seed, A.data());
  } else if (matrix == "gmm") {
    m = d;
    A.resize((size_t)d * n);
    st = rbrp_matrix_gmm_h(n, d, seed, A.data());
  } else {
    std::fprintf(stderr, "rbrp_driver: unknown matrix '%s'\n", matrix.c_str());
    return 2;
  }
  if (st) return fail("matrix generation", st);
  std::cout << "Create test matrix: " << seconds_since(t0) << " s\n";

  std::cout.precision(3);
  std::cout << "\nInputs:"
            << "\n---------------------"
            << "\nMatrix size:\t" << m << " x " << n << std::endl
            << "block size:\t" << bs << std::endl
            << "# blocks:\t" << nb << std::endl
            << "---------------------\n"
            << std::endl;
  std::cout << std::scientific;

  double anrm2 = 0.0;
  for (double x : A) anrm2 += x * x;
  double* dA = nullptr;
  if (hipMalloc(&dA, sizeof(double) * A.size()) != hipSuccess) return fail("hipMalloc", -1);
  if (hipMemcpy(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail("hipMemcpy", -1);

  // ---- RBRP (driver.cpp:170-187) --------------------------------------------
  std::vector<int> ranks;
  std::vector<double> error, t_ranks;
  double t_rbrp = 0.0, flops = 0.0;
  RBRP(m, n, dA, bs, ranks, error, t_ranks, t_rbrp, flops);
  if ((st = RBRP_last_status())) return fail("RBRP", st);
  std::cout << "RBRP total time: " << t_rbrp << std::endl;
  // the first nb blocks (fewer if the ID finished sooner)
  const size_t nk = std::min(ranks.size(), (size_t)nb);
  ranks.resize(nk);
  error.resize(nk);
  t_ranks.resize(nk);
  print_list(ranks, "ranks");
  print_list(error, "error");
  print_list(t_ranks, "time");

  // ---- full-matrix CPQR ID (driver.cpp:197-205) ------------------------------
  {
    std::cout << "\n[CPQR]\n";
    double t_cpqr = 0.0;
    CPQR(m, n, dA, ranks, error, t_ranks, t_cpqr, flops);
    if ((st = RBRP_last_status())) return fail("CPQR", st);
    std::cout << "CPQR total time: " << t_cpqr << std::endl;
    print_list(error, "error");
    print_list(t_ranks, "time");
  }

  // ---- RandCPQR-OS and RandLUPP-OS at the same ranks (driver.cpp:239-304) ----
  for (int which = 0; which < 2; ++which) {
    std::cout << (which == 0 ? "\n[RandCPQR-OS]\n" : "\n[RandLUPP-OS]\n");
    for (size_t i = 0; i < nk; ++i) {
      const int k = ranks[i];
      std::vector<int> sk((size_t)k), rd((size_t)(n - k));
      double* dT = nullptr;
      t0 = std::chrono::steady_clock::now();
      if (which == 0) {
        RandCPQR_OS(dA, m, n, k, sk.data(), rd.data(), dT, flops);
        st = RBRP_last_status();
      } else {
        int *dsk = nullptr, *drd = nullptr;
        RandLUPP_OS(dA, m, n, k, dsk, drd, dT, flops);
        st = RBRP_last_status();
        if (!st) st = (int)hipMemcpy(sk.data(), dsk, sizeof(int) * k, hipMemcpyDeviceToHost);
        if (!st && n > k)
          st = (int)hipMemcpy(rd.data(), drd, sizeof(int) * (n - k), hipMemcpyDeviceToHost);
        if (dsk) (void)hipFree(dsk);
        if (drd) (void)hipFree(drd);
      }
      if (!st) st = (int)hipDeviceSynchronize();
      t_ranks[i] = seconds_since(t0);
      if (st) {
        if (dT) (void)hipFree(dT);
        return fail(which == 0 ? "RandCPQR_OS" : "RandLUPP_OS", st);
      }
      std::vector<double> T((size_t)k * (n - k));
      if (n > k && hipMemcpy(T.data(), dT, sizeof(double) * T.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return fail("hipMemcpy", -1);
      if (dT) (void)hipFree(dT);
      error[i] = id_error_sq(A, m, n, sk, rd, T, anrm2);
    }
    print_list(error, "error");
    print_list(t_ranks, "time");
  }
  (void)hipFree(dA);
  return 0;
}
