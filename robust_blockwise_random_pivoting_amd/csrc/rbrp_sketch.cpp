// Sketch-based ID comparators behind the C ABI: RandCPQR_OS (gpu/randCPQR_OS.cu:20-111),
// RandLUPP_OS (gpu/randLUPP_OS.cu:14-119) and the oversampled interpolation
// matrix (gpu/interp.cu:45-105).  The sketch GEMM runs on the MFMA kernel;
// the pivoted QR is the blocked one of k_qrcp.hip (dgeqp3's organisation), the
// LU and the QR of the interpolation step are blocked onto MFMA panels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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

// Whether Y = G A (l x m times m x n) runs on the split-K 64 x 128 TN kernel
// as Y = (G^T)^T A: whole 16-row steps of the reduction, even leading
// dimensions, and enough output tiles to fill the device without splitting K.
bool sketch_on_tn64(rbrp_handle_t h, int64_t l, int64_t m, int64_t n, int64_t lda) {
  const int64_t tiles = ((l + 63) / 64) * ((n + 127) / 128);
  return m % 16 == 0 && lda % 2 == 0 && h->tn64_resident > 0 && 2 * tiles >= h->tn64_resident;
}

// Scratch bytes of sketch_gemm: G^T (m x l) when the TN kernel takes the product.
size_t sketch_scratch_bytes(rbrp_handle_t h, int64_t l, int64_t m, int64_t n, int64_t lda) {
  return sketch_on_tn64(h, l, m, n, lda) ? align256(sizeof(double) * m * l) : 0;
}

// Y (l x n, ldy = l) = G (l x m, ldg) * A (m x n, lda).  With Gt (scratch of
// sketch_scratch_bytes) the sketch is transposed once (0.7 ms at 4094 x 50000)
// and the product runs on the projection's TN kernel, written straight into Y
// (one K chunk); otherwise on the 64 x 64-tile NN kernel.
hipError_t sketch_gemm(rbrp_handle_t h, int64_t l, int64_t m, int64_t n, const double* G, int64_t ldg,
                       const double* A, int64_t lda, double* Y, double* Gt) {
  hipStream_t st = h->stream;
  if (Gt && sketch_on_tn64(h, l, m, n, lda)) {
    hipError_t e = rbrp_launch_transpose(st, l, m, G, ldg, Gt, m);
    if (e != hipSuccess) return e;
    return rbrp_launch_gemm_tn64(st, dim_static(l), dim_static(n), l, n, m, 1, m, mat(Gt, m),
                                 mat((double*)A, lda), Y, l, l * n);
  }
  return rbrp_launch_gemm_nn(st, dim_static(l), l, dim_static(n), n, dim_static(m),
                             mat((double*)G, ldg), mat((double*)A, lda), mat(Y, l), GEMM_SET,
                             nullptr, 0);
}

// The oversampled interpolation matrix (interp.cu:45-105): Yp = Y(:, P), QR of
// Yp(:, 0:k) applied to all of Yp (geqrf + ormqr with Q^T), then
// T (k x (n-k), ldt) = R(0:k, 0:k)^{-1} Yp(0:k, k:n).
//
// Blocked Householder QR on a copy of Yp whose leading dimension lp is l
// rounded up to 16 with zero pad rows: each panel of up to IP_NB columns is
// factored column by column, one launch per column (the reflector applied to
// the panel's columns only), then the trailing columns C get the block reflector at once,
// C -= V (T^T (V^T C)), with V the panel's explicit vectors (zero pad rows, so
// every product runs over a multiple of 16 rows), S = V^T V and T^T from
// k_wy_factor.  V^T C and S are split-K MFMA products; T^T W and V W2 are
// 64 x 64-tile MFMA products.
constexpr int64_t IP_NB = 32;      // panel width
constexpr int64_t IP_SPLITS = 16;  // split-K groups of V^T C at most

int64_t round16(int64_t x) { return (x + 15) / 16 * 16; }

size_t interp_work_bytes(int64_t l, int64_t n, int64_t k) {
  const int64_t lp = round16(l), nb = IP_NB;
  return align256(sizeof(double) * lp * n) + align256(sizeof(double) * lp * nb) +
         2 * align256(sizeof(double) * nb * nb) + 2 * align256(sizeof(double) * nb * n) +
         align256(sizeof(double) * k) +
         align256(sizeof(double) * std::max(IP_SPLITS * nb * n, 64 * nb * nb)) +
         align256(sizeof(double) * 4 * ((k + 255) / 256) * 4096) + align256(sizeof(double) * 4);
}

// split-K of a product over K rows (K a multiple of 16) for a (rows x cols)
// output: about 1024 workgroups, at most max_splits, chunks of >= 64 rows
void tn_split(int64_t K, int64_t tiles, int64_t max_splits, int* ns, int64_t* kc) {
  int64_t q = std::max<int64_t>(1, 1024 / std::max<int64_t>(tiles, 1));
  q = std::min(q, std::min(max_splits, std::max<int64_t>(1, K / 64)));
  int64_t c = round16((K + q - 1) / q);
  *kc = c;
  *ns = (int)std::max<int64_t>(1, (K + c - 1) / c);
}

hipError_t interp_os(hipStream_t st, int64_t l, int64_t n, int64_t k, const double* Y, int64_t ldy,
                     const int* P, double* T, int64_t ldt, void* work) {
  const int64_t lp = round16(l), nb = IP_NB;
  char* p = (char*)work;
  double* Yp = (double*)p;
  p += align256(sizeof(double) * lp * n);
  double* V = (double*)p;
  p += align256(sizeof(double) * lp * nb);
  double* S = (double*)p;
  p += align256(sizeof(double) * nb * nb);
  double* Tt = (double*)p;
  p += align256(sizeof(double) * nb * nb);
  double* W = (double*)p;
  p += align256(sizeof(double) * nb * n);
  double* W2 = (double*)p;
  p += align256(sizeof(double) * nb * n);
  double* tau = (double*)p;
  p += align256(sizeof(double) * k);
  double* Pbuf = (double*)p;
  p += align256(sizeof(double) * std::max(IP_SPLITS * nb * n, 64 * nb * nb));
  double* Dinv = (double*)p;
  p += align256(sizeof(double) * 4 * ((k + 255) / 256) * 4096);
  double* scr = (double*)p;
  hipError_t e;
#define IPCHK(x)                          \
  do {                                    \
    if ((e = (x)) != hipSuccess) return e; \
  } while (0)
  // (the calls that reach here synchronise before returning, so no graph
  // captures this memset)
  IPCHK(hipMemsetAsync(Yp, 0, sizeof(double) * lp * n, st));
  IPCHK(rbrp_launch_gather_cols(st, l, n, Y, ldy, P, Yp, lp));
  for (int64_t j0 = 0; j0 < k; j0 += nb) {
    const int64_t jb = std::min(nb, k - j0), je = j0 + jb;
    // panel steps: one launch per column (reflector and its application to the panel)
    for (int64_t j = j0; j < je; ++j) IPCHK(rbrp_launch_hh_step(st, l, je, j, Yp, lp, tau, scr, j > j0));
    IPCHK(rbrp_launch_hh_finish(st, l, je - 1, Yp, lp, scr));
    if (je >= n) continue;
    const int64_t rows = l - j0, rows_pad = lp - j0, nc = n - je;
    double* C = Yp + j0 + je * lp;
    IPCHK(rbrp_launch_panel_vectors(st, rows, rows_pad, (int)jb, Yp + j0 + j0 * lp, lp, V, rows_pad));
    int ns;
    int64_t kc;
    tn_split(rows_pad, 1, 64, &ns, &kc);
    IPCHK(rbrp_launch_gemm_tn(st, dim_static(jb), dim_static(jb), jb, jb, rows_pad, ns, kc, mat(V, rows_pad),
                              mat(V, rows_pad), Pbuf, jb, jb * jb));
    IPCHK(rbrp_launch_reduce_splits(st, dim_static(jb), dim_static(jb), jb, jb, Pbuf, jb, jb * jb, ns,
                                    mat(S, jb), GEMM_SET));
    IPCHK(rbrp_launch_wy_factor(st, (int)jb, S, jb, tau + j0, Tt, jb));
    tn_split(rows_pad, (nc + 63) / 64, IP_SPLITS, &ns, &kc);
    IPCHK(rbrp_launch_gemm_tn(st, dim_static(jb), dim_static(nc), jb, nc, rows_pad, ns, kc, mat(V, rows_pad),
                              mat(C, lp), Pbuf, jb, jb * nc));
    IPCHK(rbrp_launch_reduce_splits(st, dim_static(jb), dim_static(nc), jb, nc, Pbuf, jb, jb * nc, ns,
                                    mat(W, jb), GEMM_SET));
    IPCHK(rbrp_launch_gemm_nn(st, dim_static(jb), jb, dim_static(nc), nc, dim_static(jb), mat(Tt, jb),
                              mat(W, jb), mat(W2, jb), GEMM_SET, nullptr, 0));
    IPCHK(rbrp_launch_gemm_nn(st, dim_static(rows_pad), rows_pad, dim_static(nc), nc, dim_static(jb),
                              mat(V, rows_pad), mat(W2, jb), mat(C, lp), GEMM_SUB, nullptr, 0));
  }
  if (n > k) {
    IPCHK(rbrp_launch_copy_block(st, k, n - k, Yp + k * lp, lp, T, ldt));
    IPCHK(rbrp_trsm_upper_left(st, k, n - k, Yp, lp, T, ldt, nullptr, Dinv));
  }
#undef IPCHK
  return hipSuccess;
}

// LU with partial pivoting of Yt (n x k, ldy = n, k <= n) in place, blocked
// right-looking (dgetrf's order): each LU_NB-column panel is factored column by
// column with its row swaps confined to the panel, then the panel's swaps
// reach the other columns (k_lu_laswp), U12 = L11^{-1} A12 is formed in place
// (a one-row-tile product, so it may overwrite its own operand) and
// A22 -= L21 U12 is one MFMA product.  Then the row interchanges are applied
// in order to the identity (gpu/permute.cu:8-23); the permutation goes to P
// (device, n entries), its first k entries to d_sk and the rest to d_rd.
// Linv is LU_NB x LU_NB scratch.
constexpr int64_t LU_NB = 32;

// Blocked LU with partial pivoting of Y (rows x cols, leading dimension ld)
// in place to min(rows, cols) steps: LU_NB-column panels with the swaps
// confined to the panel, then the panel's interchanges on the other columns,
// U12 = L11^{-1} A12 and A22 -= L21 U12 (dgetrf's order).  ipiv[j] (0-based,
// device) is the row swapped with row j.
hipError_t lupp_factor(hipStream_t st, int64_t rows, int64_t cols, double* Y, int64_t ld, int* ipiv,
                       double* Linv) {
  const int64_t steps = std::min(rows, cols);
  hipError_t e;
  for (int64_t j0 = 0; j0 < steps; j0 += LU_NB) {
    const int64_t je = std::min(j0 + LU_NB, steps), jb = je - j0;
    for (int64_t j = j0; j < je; ++j) {
      if ((e = rbrp_launch_lu_pivot(st, rows, j0, je, j, Y, ld, ipiv)) != hipSuccess) return e;
      if ((e = rbrp_launch_lu_update(st, rows, je, j, Y, ld)) != hipSuccess) return e;
    }
    if ((e = rbrp_launch_lu_laswp(st, cols, j0, je, Y, ld, ipiv)) != hipSuccess) return e;
    if (je < cols) {
      double* A12 = Y + j0 + je * ld;
      if ((e = rbrp_launch_unit_lower_inv(st, (int)jb, Y + j0 + j0 * ld, ld, Linv, LU_NB)) != hipSuccess)
        return e;
      if ((e = rbrp_launch_gemm_nn(st, dim_static(jb), jb, dim_static(cols - je), cols - je, dim_static(jb),
                                   mat(Linv, LU_NB), mat(A12, ld), mat(A12, ld), GEMM_SET, nullptr, 0)) !=
          hipSuccess)
        return e;
      if (je < rows &&
          (e = rbrp_launch_gemm_nn(st, dim_static(rows - je), rows - je, dim_static(cols - je), cols - je,
                                   dim_static(jb), mat(Y + je + j0 * ld, ld), mat(A12, ld),
                                   mat(Y + je + je * ld, ld), GEMM_SUB, nullptr, 0)) != hipSuccess)
        return e;
    }
  }
  return hipSuccess;
}

hipError_t lupp_permutation(hipStream_t st, int64_t n, int64_t k, double* Yt, int* ipiv, int* P,
                            int32_t* d_sk, int32_t* d_rd, double* Linv) {
  const int64_t steps = std::min(n, k);
  hipError_t e;
  if ((e = lupp_factor(st, n, k, Yt, n, ipiv, Linv)) != hipSuccess) return e;
  std::vector<int> hpiv((size_t)steps), hP((size_t)n);
  if ((e = hipMemcpyAsync(hpiv.data(), ipiv, sizeof(int) * steps, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  for (int64_t q = 0; q < n; ++q) hP[q] = (int)q;
  for (int64_t j = 0; j < steps; ++j) std::swap(hP[j], hP[hpiv[j]]);
  if ((e = hipMemcpyAsync(P, hP.data(), sizeof(int) * n, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(d_sk, P, sizeof(int) * k, hipMemcpyDeviceToDevice, st)) != hipSuccess) return e;
  if (n > k && (e = hipMemcpyAsync(d_rd, P + k, sizeof(int) * (n - k), hipMemcpyDeviceToDevice, st)) != hipSuccess)
    return e;
  // (hP is a stack vector: wait for the upload before it goes out of scope)
  return hipStreamSynchronize(st);
}

// Column-pivoted Householder QR of W (r x n, ldw = r) in place to min(r, n)
// steps by the blocked pivoted QR; the pivot order goes to jpvt (device, n
// entries).  work: rbrp_qrcp_work_bytes(r, n) bytes.
hipError_t qrcp_in_place(rbrp_handle_t h, int64_t r, int64_t n, double* W, int* jpvt, double* tau,
                         void* work) {
  return rbrp_qrcp(h->stream, r, n, W, r, jpvt, tau, work, h->host_scalar);
}

int check_common(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda, int64_t r,
                 const double* dG, int64_t ldg, double* dT, int64_t ldt) {
  if (!h) return -1;
  if (m < 1) return -3;
  if (n < 1) return -4;
  if (lda < m) return -5;
  if (!dA) return -2;
  if (r < 1 || r > std::min(m, n)) return -6;
  if (dG && ldg < 2 * r) return -10;
  if (n > r && (!dT || ldt < r)) return -13;
  (void)dG;
  return 0;
}

}  // namespace

extern "C" {

int rbrp_gaussian_sketch_d(rbrp_handle_t h, uint64_t seed, uint64_t offset, int64_t rows,
                           int64_t cols, double scale, double* dG, int64_t ldg) {
  if (!h) return -1;
  if (rows < 0) return -4;
  if (cols < 0) return -5;
  if (rows * cols > 0 && !dG) return -7;
  if (ldg < rows) return -8;
  HIPCHK(rbrp_launch_gaussian(h->stream, seed, offset, rows, cols, scale, dG, ldg));
  return 0;
}

int rbrp_interp_os_d(rbrp_handle_t h, const double* dY, int64_t l, int64_t n, int64_t ldy, int64_t k,
                     const int32_t* dP, double* dT, int64_t ldt) {
  if (!h) return -1;
  if (!dY) return -2;
  if (l < 1) return -3;
  if (n < 1) return -4;
  if (ldy < l) return -5;
  if (k < 1 || k > std::min(l, n)) return -6;
  if (!dP) return -7;
  if (n > k && (!dT || ldt < k)) return -8;
  hipStream_t st = h->stream;
  void* arena = nullptr;
  HIPCHK(rbrp_sketch_arena(h, interp_work_bytes(l, n, k), &arena));
  HIPCHK(interp_os(st, l, n, k, dY, ldy, dP, dT, ldt, arena));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int rbrp_randcpqr_os_d(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda,
                       int64_t r, uint64_t seed, uint64_t offset, const double* dG, int64_t ldg,
                       int32_t* h_sk, int32_t* h_rd, double* dT, int64_t ldt) {
  int c = check_common(h, dA, m, n, lda, r, dG, ldg, dT, ldt);
  if (c) return c;
  if (!h_sk) return -11;
  if (n > r && !h_rd) return -12;
  hipStream_t st = h->stream;
  const int64_t l = 2 * r;  // oversampling (randCPQR_OS.cu:24)
  const size_t bG = dG ? 0 : align256(sizeof(double) * l * m);
  const size_t bY = align256(sizeof(double) * l * n), bW = align256(sizeof(double) * r * n);
  const size_t bvn = align256(rbrp_qrcp_work_bytes(r, n)), btau = align256(sizeof(double) * l);
  const size_t bjp = align256(sizeof(int) * n);
  void* arena = nullptr;
  const size_t bGt = sketch_scratch_bytes(h, l, m, n, lda);
  const size_t base = bG + bY + bW + bvn + btau + bjp + interp_work_bytes(l, n, r);
  HIPCHK(rbrp_sketch_arena(h, base + bGt, &arena));
  double* Gt = bGt ? (double*)((char*)arena + base) : nullptr;
  char* p = (char*)arena;
  double* G = dG ? (double*)dG : (double*)p;
  const int64_t ldG = dG ? ldg : l;
  p += bG;
  double* Y = (double*)p;
  p += bY;
  double* W = (double*)p;
  p += bW;
  void* qwork = p;
  p += bvn;
  double* tau = (double*)p;
  p += btau;
  int* jpvt = (int*)p;
  p += bjp;
  void* iwork = p;
  if (!dG) HIPCHK(rbrp_launch_gaussian(st, seed, offset, l, m, 1.0 / std::sqrt((double)m), G, l));
  HIPCHK(sketch_gemm(h, l, m, n, G, ldG, dA, lda, Y, Gt));      // randCPQR_OS.cu:42
  // pivoted QR of Y(0:r, :) (randCPQR_OS.cu:55-72): only the pivots are used
  HIPCHK(rbrp_launch_copy_block(st, r, n, Y, l, W, r));
  HIPCHK(qrcp_in_place(h, r, n, W, jpvt, tau, qwork));
  // T by the oversampled interpolation (randCPQR_OS.cu:95, interp.cu:45-105)
  HIPCHK(interp_os(st, l, n, r, Y, l, jpvt, dT, ldt, iwork));
  std::vector<int32_t> hp((size_t)n);
  HIPCHK(hipMemcpyAsync(hp.data(), jpvt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::copy(hp.begin(), hp.begin() + r, h_sk);
  if (n > r) std::copy(hp.begin() + r, hp.end(), h_rd);
  return 0;
}

int rbrp_randlupp_os_d(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda,
                       int64_t k, uint64_t seed, uint64_t offset, const double* dG, int64_t ldg,
                       int32_t* d_sk, int32_t* d_rd, double* dT, int64_t ldt) {
  int c = check_common(h, dA, m, n, lda, k, dG, ldg, dT, ldt);
  if (c) return c;
  if (!d_sk) return -11;
  if (n > k && !d_rd) return -12;
  hipStream_t st = h->stream;
  const int64_t l = 2 * k;  // oversampling (randLUPP_OS.cu:18)
  const size_t bG = dG ? 0 : align256(sizeof(double) * l * m);
  const size_t bY = align256(sizeof(double) * l * n), bYt = align256(sizeof(double) * n * k);
  const size_t bip = align256(sizeof(int) * k);
  const size_t bP = align256(sizeof(int) * n);
  void* arena = nullptr;
  const size_t bLinv = align256(sizeof(double) * LU_NB * LU_NB);
  const size_t bGt = sketch_scratch_bytes(h, l, m, n, lda);
  const size_t base = bG + bY + bYt + bip + bP + bLinv + interp_work_bytes(l, n, k);
  HIPCHK(rbrp_sketch_arena(h, base + bGt, &arena));
  double* Gt = bGt ? (double*)((char*)arena + base) : nullptr;
  char* p = (char*)arena;
  double* G = dG ? (double*)dG : (double*)p;
  const int64_t ldG = dG ? ldg : l;
  p += bG;
  double* Y = (double*)p;
  p += bY;
  double* Yt = (double*)p;
  p += bYt;
  int* ipiv = (int*)p;
  p += bip;
  int* P = (int*)p;
  p += bP;
  double* Linv = (double*)p;
  p += bLinv;
  void* iwork = p;
  if (!dG) HIPCHK(rbrp_launch_gaussian(st, seed, offset, l, m, 1.0 / std::sqrt((double)m), G, l));
  HIPCHK(sketch_gemm(h, l, m, n, G, ldG, dA, lda, Y, Gt));      // randLUPP_OS.cu:43
  // LU with partial pivoting of Y(0:k, :)^T (n x k) (randLUPP_OS.cu:52-82)
  HIPCHK(rbrp_launch_transpose(st, k, n, Y, l, Yt, n));
  // pivots -> permutation (randLUPP_OS.cu:85, permute.cu:8-23)
  HIPCHK(lupp_permutation(st, n, k, Yt, ipiv, P, d_sk, d_rd, Linv));
  // T by the oversampled interpolation (randLUPP_OS.cu:96, interp.cu:45-105)
  HIPCHK(interp_os(st, l, n, k, Y, l, P, dT, ldt, iwork));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int rbrp_randcpqr_d(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda, int64_t r,
                    uint64_t seed, uint64_t offset, const double* dG, int64_t ldg, int32_t* h_sk,
                    int32_t* h_rd, double* dT, int64_t ldt) {
  int c = check_common(h, dA, m, n, lda, r, nullptr, 0, dT, ldt);
  if (c) return c;
  if (dG && ldg < r) return -10;
  if (!h_sk) return -11;
  if (n > r && !h_rd) return -12;
  hipStream_t st = h->stream;
  const size_t bG = dG ? 0 : align256(sizeof(double) * r * m);
  const size_t bW = align256(sizeof(double) * r * n), bvn = align256(rbrp_qrcp_work_bytes(r, n));
  const size_t btau = align256(sizeof(double) * r), bjp = align256(sizeof(int) * n);
  const size_t bDinv = align256(sizeof(double) * 4 * ((r + 255) / 256) * 4096);
  void* arena = nullptr;
  const size_t bGt = sketch_scratch_bytes(h, r, m, n, lda);
  const size_t base = bG + bW + bvn + btau + bjp + bDinv;
  HIPCHK(rbrp_sketch_arena(h, base + bGt, &arena));
  double* Gt = bGt ? (double*)((char*)arena + base) : nullptr;
  char* p = (char*)arena;
  double* G = dG ? (double*)dG : (double*)p;
  const int64_t ldG = dG ? ldg : r;
  p += bG;
  double* W = (double*)p;
  p += bW;
  void* qwork = p;
  p += bvn;
  double* tau = (double*)p;
  p += btau;
  int* jpvt = (int*)p;
  p += bjp;
  double* Dinv = (double*)p;
  if (!dG) HIPCHK(rbrp_launch_gaussian(st, seed, offset, r, m, 1.0 / std::sqrt((double)m), G, r));
  HIPCHK(sketch_gemm(h, r, m, n, G, ldG, dA, lda, W, Gt));      // randCPQR.cu:42
  HIPCHK(qrcp_in_place(h, r, n, W, jpvt, tau, qwork));            // randCPQR.cu:52
  if (n > r) {
    // T = R11^{-1} R12 straight from the sketch's factor (randCPQR.cu:60-64)
    HIPCHK(rbrp_launch_copy_block(st, r, n - r, W + r * r, r, dT, ldt));
    HIPCHK(rbrp_trsm_upper_left(st, r, n - r, W, r, dT, ldt, nullptr, Dinv));
  }
  std::vector<int32_t> hp((size_t)n);
  HIPCHK(hipMemcpyAsync(hp.data(), jpvt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::copy(hp.begin(), hp.begin() + r, h_sk);
  if (n > r) std::copy(hp.begin() + r, hp.end(), h_rd);
  return 0;
}

int rbrp_randlupp_d(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda, int64_t k,
                    uint64_t seed, uint64_t offset, const double* dG, int64_t ldg, int32_t* d_sk,
                    int32_t* d_rd, double* dT, int64_t ldt) {
  int c = check_common(h, dA, m, n, lda, k, nullptr, 0, dT, ldt);
  if (c) return c;
  if (dG && ldg < k) return -10;
  if (!d_sk) return -11;
  if (n > k && !d_rd) return -12;
  hipStream_t st = h->stream;
  const size_t bG = dG ? 0 : align256(sizeof(double) * k * m);
  const size_t bY = align256(sizeof(double) * k * n), bYt = bY, bL = align256(sizeof(double) * k * k);
  const size_t bip = align256(sizeof(int) * k), bP = align256(sizeof(int) * n);
  const size_t bDinv = align256(sizeof(double) * 4 * ((k + 255) / 256) * 4096);
  void* arena = nullptr;
  const size_t bLinv = align256(sizeof(double) * LU_NB * LU_NB);
  const size_t bGt = sketch_scratch_bytes(h, k, m, n, lda);
  const size_t base = bG + bY + bYt + bL + bip + bP + bDinv + bLinv;
  HIPCHK(rbrp_sketch_arena(h, base + bGt, &arena));
  double* Gt = bGt ? (double*)((char*)arena + base) : nullptr;
  char* p = (char*)arena;
  double* G = dG ? (double*)dG : (double*)p;
  const int64_t ldG = dG ? ldg : k;
  p += bG;
  double* Y = (double*)p;
  p += bY;
  double* Yt = (double*)p;
  p += bYt;
  double* Lt = (double*)p;
  p += bL;
  int* ipiv = (int*)p;
  p += bip;
  int* P = (int*)p;
  p += bP;
  double* Dinv = (double*)p;
  p += bDinv;
  double* Linv = (double*)p;
  if (!dG) HIPCHK(rbrp_launch_gaussian(st, seed, offset, k, m, 1.0 / std::sqrt((double)m), G, k));
  HIPCHK(sketch_gemm(h, k, m, n, G, ldG, dA, lda, Y, Gt));      // randLUPP.cu:43
  // LU with partial pivoting of Y^T (n x k) (randLUPP.cu:47-78) and its permutation (:82-85)
  HIPCHK(rbrp_launch_transpose(st, k, n, Y, k, Yt, n));
  HIPCHK(lupp_permutation(st, n, k, Yt, ipiv, P, d_sk, d_rd, Linv));
  if (n > k) {
    // T = (L21 L11^{-1})^T = L11^{-T} L21^T (randLUPP.cu:89-106), solved as an
    // upper unit-triangular system with L11^T
    HIPCHK(rbrp_launch_transpose(st, k, k, Yt, n, Lt, k));
    HIPCHK(rbrp_launch_unit_upper(st, k, Lt, k));
    HIPCHK(rbrp_launch_transpose(st, n - k, k, Yt + k, n, dT, ldt));
    HIPCHK(rbrp_trsm_upper_left(st, k, n - k, Lt, k, dT, ldt, nullptr, Dinv));
  }
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int rbrp_randadaplupp_d(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda,
                        double tol, int64_t blk, uint64_t seed, uint64_t offset, int32_t* d_sk,
                        int32_t* d_rd, double* dT, int64_t ldt, int64_t* rank, double* err_est,
                        double* flops) {
  if (!h) return -1;
  if (m < 1) return -3;
  if (n < 1) return -4;
  if (lda < m) return -5;
  if (!dA) return -2;
  if (!(tol >= 0.0) || !std::isfinite(tol)) return -6;
  if (blk < 1 || blk > LU_NB) return -7;
  if (!d_sk) return -10;
  if (!d_rd) return -11;
  const int64_t p = std::min(m, n);
  if (!dT) return -12;
  if (ldt < p) return -13;
  if (!rank) return -14;
  hipStream_t st = h->stream;
  const int64_t b = blk;
  // workspace: sketch G (b x m), Y = G A (b x n), the transposed residual Rt
  // (n x b, rows in permutation order), the skeleton part Ys (p x b), the
  // interpolation matrix M (n x p, rows in permutation order), P, pivots, the
  // unit-lower inverse of a panel, column norms and their sum
  const size_t bG = align256(sizeof(double) * b * m), bY = align256(sizeof(double) * b * n);
  const size_t bRt = align256(sizeof(double) * n * b), bYs = align256(sizeof(double) * p * b);
  const size_t bM = align256(sizeof(double) * n * p), bP = align256(sizeof(int) * n);
  const size_t bip = align256(sizeof(int) * b), bLinv = align256(sizeof(double) * LU_NB * LU_NB);
  const size_t bcn = align256(sizeof(double) * (std::max(n, b) + 1));
  const size_t bGt = sketch_scratch_bytes(h, b, m, n, lda);
  const size_t base = bG + bY + bRt + bYs + bM + bP + bip + bLinv + bcn;
  void* arena = nullptr;
  HIPCHK(rbrp_sketch_arena(h, base + bGt, &arena));
  char* q = (char*)arena;
  double* G = (double*)q;
  q += bG;
  double* Y = (double*)q;
  q += bY;
  double* Rt = (double*)q;
  q += bRt;
  double* Ys = (double*)q;
  q += bYs;
  double* M = (double*)q;
  q += bM;
  int* P = (int*)q;
  q += bP;
  int* ipiv = (int*)q;
  q += bip;
  double* Linv = (double*)q;
  q += bLinv;
  double* cn = (double*)q;
  double* Gt = bGt ? (double*)((char*)arena + base) : nullptr;
  double fl = 0.0;
  auto sumsq = [&](const double* X, int64_t rows, int64_t cols, int64_t ld, double* out) -> int {
    HIPCHK(rbrp_launch_colnorm(st, X, ld, rows, cols, cn));
    HIPCHK(rbrp_launch_sum(st, cn, cols, cn + std::max(n, b)));
    HIPCHK(hipMemcpyAsync(out, cn + std::max(n, b), sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  };
  // block t's sketch: G_t ~ N(0, 1/b) from the Philox stream (seed, offset + t b m)
  auto sketch = [&](int64_t t) -> int {
    HIPCHK(rbrp_launch_gaussian(st, seed, offset + (uint64_t)t * (uint64_t)(b * m), b, m,
                                1.0 / std::sqrt((double)b), G, b));
    HIPCHK(sketch_gemm(h, b, m, n, G, b, dA, lda, Y, Gt));
    fl += 2.0 * b * m * n;
    return 0;
  };
  double normA2 = 0.0;
  {
    const int e = sumsq(dA, m, n, lda, &normA2);
    if (e) return e;
  }
  const double tol2 = tol * tol * normA2;
  HIPCHK(rbrp_launch_iota(st, P, n));
  // block 0: the residual is the first sketch itself
  {
    const int e = sketch(0);
    if (e) return e;
  }
  HIPCHK(rbrp_launch_gather_cols_t(st, n, b, Y, b, P, 0, Rt, n));
  int64_t k = 0, t = 0;
  double est2 = normA2;
  while (k < p) {
    // (a block never takes the rank past min(m, n), the width of M)
    const int64_t bt = std::min(b, p - k), nc = n - k;
    // LU with partial pivoting of the candidates' residual (nc x bt)
    HIPCHK(lupp_factor(st, nc, bt, Rt + k, n, ipiv, Linv));
    fl += (double)nc * bt * bt;
    // its interchanges on the candidates' rows of M and on P
    HIPCHK(rbrp_launch_swap_rows(st, (int)bt, k, ipiv, M, n, k, P));
    // M_b = L2b Ltt^{-1} for the new columns; the old columns M(rest) -= M_b M(new)
    if (nc > bt) {
      HIPCHK(rbrp_launch_unit_lower_inv(st, (int)bt, Rt + k, n, Linv, LU_NB));
      HIPCHK(rbrp_launch_gemm_nn(st, dim_static(nc - bt), nc - bt, dim_static(bt), bt, dim_static(bt),
                                 mat(Rt + k + bt, n), mat(Linv, LU_NB), mat(M + (k + bt) + k * n, n),
                                 GEMM_SET, nullptr, 0));
      if (k > 0)
        HIPCHK(rbrp_launch_gemm_nn(st, dim_static(nc - bt), nc - bt, dim_static(k), k, dim_static(bt),
                                   mat(M + (k + bt) + k * n, n), mat(M + k, n), mat(M + (k + bt), n),
                                   GEMM_SUB, nullptr, 0));
      fl += 2.0 * (nc - bt) * bt * (bt + k);
    }
    k += bt;
    ++t;
    if (k >= p || k >= n) {
      est2 = 0.0;
      break;
    }
    // a fresh sketch: the estimate of the ID error and the next block's residual,
    // Rt(k:n) = Y(:, P[k:n])^T - M(k:n, 0:k) Y(:, P[0:k])^T
    {
      const int e = sketch(t);
      if (e) return e;
    }
    HIPCHK(rbrp_launch_gather_cols_t(st, k, b, Y, b, P, 0, Ys, p));
    HIPCHK(rbrp_launch_gather_cols_t(st, n - k, b, Y, b, P, k, Rt + k, n));
    HIPCHK(rbrp_launch_gemm_nn(st, dim_static(n - k), n - k, dim_static(b), b, dim_static(k),
                               mat(M + k, n), mat(Ys, p), mat(Rt + k, n), GEMM_SUB, nullptr, 0));
    fl += 2.0 * (n - k) * k * b;
    {
      const int e = sumsq(Rt + k, n - k, b, n, &est2);
      if (e) return e;
    }
    if (est2 <= tol2) break;
  }
  // outputs: sk = P[0:k], rd = P[k:n], T = M(k:n, 0:k)^T (k x (n-k))
  HIPCHK(hipMemcpyAsync(d_sk, P, sizeof(int) * k, hipMemcpyDeviceToDevice, st));
  if (n > k) {
    HIPCHK(hipMemcpyAsync(d_rd, P + k, sizeof(int) * (n - k), hipMemcpyDeviceToDevice, st));
    HIPCHK(rbrp_launch_transpose(st, n - k, k, M + k, n, dT, ldt));
  }
  HIPCHK(hipStreamSynchronize(st));
  *rank = k;
  if (err_est) *err_est = normA2 > 0.0 ? std::sqrt(std::max(est2, 0.0) / normA2) : 0.0;
  if (flops) *flops = fl;
  return 0;
}

}  // extern "C"
