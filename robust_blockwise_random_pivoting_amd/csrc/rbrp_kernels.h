// Kernel launchers of the RBRP HIP library.  Each kernel translation unit
// exports plain host functions that launch its kernels on a given stream, so
// the host orchestration never names a __global__ symbol directly.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "rbrp_common.h"

// Largest block size b of the column ID.
#define RBRP_MAX_B 128

// Kernel launch that, given start / stop events (either may be null), has the
// dispatch itself time them (hipExtLaunchKernelGGL): no marker packets enter
// the stream around the kernel.  The kernel name may be parenthesised.
#define RBRP_LAUNCH_EV(e0, e1, kern, grid, block, shm, st, ...)                    \
  do {                                                                            \
    if ((e0) != nullptr || (e1) != nullptr)                                       \
      hipExtLaunchKernelGGL(kern, grid, block, shm, st, e0, e1, 0, __VA_ARGS__);  \
    else                                                                          \
      hipLaunchKernelGGL(kern, grid, block, shm, st, __VA_ARGS__);                \
  } while (0)

// hipFuncSetAttribute(fn, MaxDynamicSharedMemorySize, bytes), called only when
// the largest size set so far for this function on the current device is
// smaller.  `cache` is the call site's zero-initialised static array (one slot
// per device); concurrent callers at worst set the same value twice.
#define RBRP_MAX_DEVICES 64
static inline hipError_t rbrp_allow_lds(const void* fn, int bytes, std::atomic<int>* cache) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= RBRP_MAX_DEVICES)
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (cache[dev].load(std::memory_order_relaxed) >= bytes) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess) cache[dev].store(bytes, std::memory_order_relaxed);
  return e;
}

// A size that is either static (`base`) or `base + sign * (*dyn)` read on the
// device, so one launch shape serves every block of the loop.
struct DimSpec {
  int64_t base;
  const int* dyn;
  int sign;
};

static inline DimSpec dim_static(int64_t v) { return DimSpec{v, nullptr, 0}; }
static inline DimSpec dim_dyn(int64_t base, const int* dyn, int sign) {
  return DimSpec{base, dyn, sign};
}

// Column-major operand: column c lives at p + col(c) * ld, where col(c) is
// cols[c] when an index list is given, else coff + (*coff_dyn) + c.
struct MatRef {
  double* p;
  int64_t ld;
  const int* cols;
  int64_t coff;
  const int* coff_dyn;
};

static inline MatRef mat(double* p, int64_t ld) { return MatRef{p, ld, nullptr, 0, nullptr}; }
static inline MatRef mat_off(double* p, int64_t ld, int64_t coff, const int* coff_dyn = nullptr) {
  return MatRef{p, ld, nullptr, coff, coff_dyn};
}
static inline MatRef mat_cols(double* p, int64_t ld, const int* cols) {
  return MatRef{p, ld, cols, 0, nullptr};
}

#if defined(__HIPCC__)
__device__ inline int64_t dim_eval(const DimSpec& d) {
  return d.dyn ? d.base + (int64_t)d.sign * (int64_t)(*d.dyn) : d.base;
}
__device__ inline double* mat_col(const MatRef& r, int64_t c) {
  const int64_t cc = r.cols ? (int64_t)r.cols[c] : (r.coff + (r.coff_dyn ? (int64_t)(*r.coff_dyn) : 0) + c);
  return r.p + cc * r.ld;
}
#endif

enum { GEMM_SET = 0, GEMM_SUB = 1, GEMM_ADD = 2 };

// The caller's matrix as the source of the FP64 loop's first block (before a
// working copy exists): column c is p + cols[coff + *coff_dyn + c] * ld (the
// permuted order), rows >= mtrue read as zero.
struct FirstSrc {
  const double* p;
  int64_t ld;
  const int* cols;
  int64_t coff;
  const int* coff_dyn;
  int64_t mtrue;
};

// ---- k_gemm.hip -----------------------------------------------------------
// Split-K TN product: P[z](r, c) = sum_{kk in chunk z} A(kk, r) * B(kk, c),
// r < M, c < N.  K must be a multiple of 16 and rows kk >= true height must be
// zero (the padded buffers guarantee it).  Grid: ceil(Nmax/64) x ceil(Mmax/64) x nsplit.
hipError_t rbrp_launch_gemm_tn(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                               int64_t K, int nsplit, int64_t kchunk, MatRef A, MatRef B,
                               double* P, int64_t ldp, int64_t pstride);

// The same product on 64 x 128 output tiles with double-buffered LDS (the
// projection and the re-orthogonalisation coefficients).  K and kchunk
// multiples of 16, both leading dimensions even.
// With `first`, B is read from the caller's matrix instead (B is then unused).
hipError_t rbrp_launch_gemm_tn64(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                                 int64_t K, int nsplit, int64_t kchunk, MatRef A, MatRef B, double* P,
                                 int64_t ldp, int64_t pstride, hipEvent_t e0 = nullptr,
                                 hipEvent_t e1 = nullptr, const FirstSrc* first = nullptr);
// Split-K Gram matrix of a tall block of at most 128 columns (the panel's
// V^T V and Q1^T Q1): P[z](r, c) = sum_{kk in chunk z} A(kk, r) A(kk, c) for
// r, c < N <= Nmax <= 128, chunks of kchunk rows (a multiple of 16, as K), the
// whole symmetric matrix written; A's leading dimension even, its columns
// 16-byte aligned.
hipError_t rbrp_launch_gram(hipStream_t st, DimSpec N, int64_t Nmax, int64_t K, int nsplit,
                            int64_t kchunk, MatRef A, double* P, int64_t ldp, int64_t pstride);
// Workgroups of k_gemm_tn64 / k_gemm_tn32 resident at once on the current
// device (compute units x occupancy); 0 if the runtime cannot tell.  `wide`:
// the FP32 kernel's 128-row tile form (launched when A is wider than 64).
int rbrp_gemm_tn64_resident();
int rbrp_gemm_tn32_resident(int wide = 0);

// C(r, c) (op)= sum_{k < K} A(r, k) B(k, c) for a tall C with 64-column tiles
// and row ranges of at most 64 rows balanced over `resident` workgroups (the
// kernel's whole-device residency, rbrp_gemm_nn64_resident; 0: about 1024),
// double-buffered K steps (the re-orthogonalisation V -= Q C, the new block
// Q_b = Q1 R2^{-1} and Q1 = V(:, Jk) Rb^{-1}); any K (a device size allowed);
// any MatRef for A.  With C32 the new C(r, c) is also stored in FP32 at
// C32[r + c * ld32].
hipError_t rbrp_launch_gemm_nn64(hipStream_t st, DimSpec M, int64_t Mmax, DimSpec N, int64_t Nmax,
                                 DimSpec K, MatRef A, MatRef B, MatRef C, int mode,
                                 int resident = 0, float* C32 = nullptr, int64_t ld32 = 0);
int rbrp_gemm_nn64_resident();
// Row ranges of k_gemm_nn64 and its norms form for an Mmax x Nmax output.
int64_t rbrp_gemm_nn64_ranges(int64_t Mmax, int64_t Nmax, int resident);
// Norms form: normpart[g * ldnp + c] = sum over row range g of
// (C(r, c) - sum_k A(r, k) B(k, c))^2, r < M (true height), c < N, for g <
// rbrp_gemm_nn64_ranges(M, N, resident); E is not stored.  With a32, A.p and
// C.p hold FP32 matrices (float pointers), widened on load.
hipError_t rbrp_launch_gemm_nn64_resnorm(hipStream_t st, int64_t M, int64_t N, int64_t K, MatRef A,
                                         MatRef B, MatRef C, int a32, double* normpart, int64_t ldnp,
                                         int resident = 0);

// dst(r, c) (=|+=|-=) sum_z P[z](r, c) in a fixed order (skipped when dst.p is
// null); when dst2 != nullptr the sum is also stored row-major at
// dst2[r * ld2 + c] (the transpose, for a column-major dst2); when radd !=
// nullptr it is also added to radd[c + racols[r] * ldra].
hipError_t rbrp_launch_reduce_splits(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                                     const double* P, int64_t ldp, int64_t pstride, int nsplit,
                                     MatRef dst, int mode, double* dst2 = nullptr, int64_t ld2 = 0,
                                     double* radd = nullptr, int64_t ldra = 0,
                                     const int* racols = nullptr);

// As rbrp_launch_reduce_splits, with the row-major copy stored in FP32.
hipError_t rbrp_launch_reduce_splits_f32copy(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax,
                                             int64_t Nmax, const double* P, int64_t ldp,
                                             int64_t pstride, int nsplit, MatRef dst, int mode,
                                             float* dst2, int64_t ld2);

// NN product with update: C(r, c) = A(r,:) B(:,c) (GEMM_SET), C -= A B (GEMM_SUB) or
// C += A B (GEMM_ADD), rows r < M, columns c < N, inner dimension K (any size),
// grid sized by Mmax x Nmax.  If normpart != nullptr, writes
// normpart[tile_row * ldnp + c] = sum over the 64-row tile of C_new(r, c)^2.
hipError_t rbrp_launch_gemm_nn(hipStream_t st, DimSpec M, int64_t Mmax, DimSpec N, int64_t Nmax,
                               DimSpec K, MatRef A, MatRef B, MatRef C, int mode, double* normpart,
                               int64_t ldnp);

// Trailing update C(r, c) -= sum_k A(r, k) Bt[k * ldbt + c] for r < M, c < N,
// K <= Kmax <= 128.  A and C are column-major with leading dimensions lda /
// ldc whose first column is acoff + *acoff_dyn / ccoff + *ccoff_dyn (the
// dynamic parts optional); Bt is row-major.  Optional squared column norms of
// the new C as partial sums over row groups: normpart[g * ldnp + c], g <
// *groups (at most ceil(M / 64) groups).
hipError_t rbrp_launch_gemm_trail(hipStream_t st, int64_t M, DimSpec N, int64_t Nmax, DimSpec K,
                                  int64_t Kmax, const double* A, int64_t lda, int64_t acoff,
                                  const int* acoff_dyn, const double* Bt, int64_t ldbt, double* C,
                                  int64_t ldc, int64_t ccoff, const int* ccoff_dyn,
                                  double* normpart, int64_t ldnp, int64_t* groups = nullptr,
                                  hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                                  const FirstSrc* first = nullptr);
// Whether rbrp_launch_gemm_trail takes a `first` source for these sizes (the
// 128-column kernel: K <= 64, M a multiple of 64, lda even).
bool rbrp_gemm_trail_first_ok(int64_t M, int64_t Kmax, int64_t lda);

// ---- k_rbrp.hip -----------------------------------------------------------
hipError_t rbrp_launch_copy_colnorm(hipStream_t st, const double* A, int64_t lda, int64_t m,
                                    int64_t mp, int64_t n, double* Aw, double* diags);
// diags[j] = ||A(:, j)||^2 (m rows), a read of A only.
hipError_t rbrp_launch_colnorm(hipStream_t st, const double* A, int64_t lda, int64_t m, int64_t n,
                               double* diags);
// V(:, p) = [A(0:m, J[p]); zero rows up to mp], p < bs->s (grid sized by b).
hipError_t rbrp_launch_gather_first(hipStream_t st, const double* A, int64_t lda, int64_t m,
                                    int64_t mp, int b, const RbrpBlockState* bs, const int* J,
                                    double* V, int64_t ldv);
hipError_t rbrp_launch_sum_to_state(hipStream_t st, const double* x, int64_t n, RbrpBlockState* bs,
                                    int which);
hipError_t rbrp_launch_iota(hipStream_t st, int* perm, int64_t n);
// *out = sum of x[0:n) in a fixed order (one workgroup).
hipError_t rbrp_launch_sum(hipStream_t st, const double* x, int64_t n, double* out);
// p[0:count) = 0 by a fill kernel.
hipError_t rbrp_launch_zero(hipStream_t st, double* p, int64_t count);
// Sampler of a block whose first column is i (+ *i_dyn with i_dyn: then the
// state holds the previous block's results and the call does nothing, with
// s = 0, when that block ended the loop; tol2 is the squared tolerance).
hipError_t rbrp_launch_sample(hipStream_t st, const double* diags, int64_t i, int64_t n,
                              int64_t kmax, int b, uint64_t key0, uint64_t key1,
                              const uint64_t ctr[4], double* w, double* cdf, RbrpBlockState* bs,
                              int* J, const int* i_dyn = nullptr, double tol2 = 0.0);
hipError_t rbrp_launch_pivchol_filter(hipStream_t st, const double* G, int64_t ldg, int b,
                                      int filter_mode, int64_t i, const int* J, RbrpBlockState* bs,
                                      double* Rb, double* Rinv, int* Jk, int* Jin, int* Jout,
                                      int* Jkpos = nullptr);
// G2 = R2^T R2: writes R2, R2inv = R2^{-1} (ld b) and Rout = R2 * Rb (ldr).
hipError_t rbrp_launch_chol2(hipStream_t st, const double* G2, int64_t ldg, int b,
                             RbrpBlockState* bs, const double* Rb, double* R2, double* R2inv,
                             double* Rout, int64_t ldr);
// One-wave forms of the two launches above (b <= 64, k_panel.hip).
hipError_t rbrp_launch_pivchol_wave(hipStream_t st, const double* G, int64_t ldg, int b,
                                    int filter_mode, int64_t i, const int* J, RbrpBlockState* bs,
                                    double* Rb, double* Rinv, int* Jk, int* Jin, int* Jout,
                                    int* Jkpos);
hipError_t rbrp_launch_chol2_wave(hipStream_t st, const double* G2, int64_t ldg, int b,
                                  RbrpBlockState* bs, const double* Rb, double* R2, double* R2inv,
                                  double* Rout, int64_t ldr);
// Workgroup forms (b <= 128, k_panel.hip): the Schur complement in registers.
hipError_t rbrp_launch_pivchol_regs(hipStream_t st, const double* G, int64_t ldg, int b,
                                    int filter_mode, int64_t i, const int* J, RbrpBlockState* bs,
                                    double* Rb, double* Rinv, int* Jk, int* Jin, int* Jout,
                                    int* Jkpos);
hipError_t rbrp_launch_chol2_regs(hipStream_t st, const double* G2, int64_t ldg, int b,
                                  RbrpBlockState* bs, const double* Rb, double* R2, double* R2inv,
                                  double* Rout, int64_t ldr);
// Inverses of the diagonal blocks of the k x k upper-triangular R in the
// partition of the blocked back substitution (256-row super-blocks from the
// bottom, 64-row blocks within): 4 * ceil(k / 256) blocks of 64 x 64 doubles
// (ld 64) at out.
hipError_t rbrp_launch_upper_inv_blocks(hipStream_t st, int64_t k, const double* R, int64_t ldr,
                                        double* out);
// Symmetric-difference permutation of one block in one launch: perm and diags
// (Jin <- Jout, the front <- Jk), the same column move in R(0:i, :), and
// Aw(:, Jin) = Aw(:, Jout) over `rows` rows (Aw null: no working-matrix move).
// b <= RBRP_MAX_B.
hipError_t rbrp_launch_perm_block(hipStream_t st, int64_t i, int b, const RbrpBlockState* bs,
                                  const int* Jk, const int* Jin, const int* Jout, int* perm,
                                  double* diags, double* R, int64_t ldr, double* Aw, int64_t rows,
                                  int64_t lda);
// The same with the FP32 working matrix of the mixed form.
hipError_t rbrp_launch_perm_block32(hipStream_t st, int64_t i, int b, const RbrpBlockState* bs,
                                    const int* Jk, const int* Jin, const int* Jout, int* perm,
                                    double* diags, double* R, int64_t ldr, float* Aw, int64_t rows,
                                    int64_t lda);
// diags[c] for the remaining columns from the update's norm partials.
hipError_t rbrp_launch_colnorm_reduce(hipStream_t st, const double* part, int64_t ldp, int64_t ntiles,
                                      int64_t i, int64_t n, RbrpBlockState* bs, double* diags);
// bs->resid / err after the block; with pub (pinned host-coherent memory) the
// state is also published there, seq written last with a system-scope release;
// with snap the same record (seq included) is also stored there on the device.
hipError_t rbrp_launch_resid_total(hipStream_t st, const double* diags, int64_t i, int64_t n,
                                   RbrpBlockState* bs, RbrpBlockState* pub = nullptr, int seq = 0,
                                   RbrpBlockState* snap = nullptr);
hipError_t rbrp_launch_copy_block(hipStream_t st, int64_t rows, int64_t cols, const double* src,
                                  int64_t lds, double* dst, int64_t ldd);
hipError_t rbrp_launch_trsm_diag(hipStream_t st, int64_t nb, int64_t ncols, const double* Rdiag,
                                 int64_t ldr, double* X, int64_t ldx);
// X (k x nc, ldx) <- R^{-1} X for the k x k upper-triangular R (ldr) in one
// launch (16 columns per workgroup), with the diagonal-block inverses of
// rbrp_launch_upper_inv_blocks at dinv.
hipError_t rbrp_launch_trsm_cols(hipStream_t st, int64_t k, int64_t nc, const double* R, int64_t ldr,
                                 const double* dinv, double* X, int64_t ldx);
hipError_t rbrp_launch_uniforms(hipStream_t st, uint64_t key0, uint64_t key1, const uint64_t ctr[4],
                                uint64_t draw0, int64_t count, double* out);

// ---- k_mixed.hip (FP32-stored residual) ------------------------------------
// Aw (mp x n, FP32) = fp32 copy of A (FP64, or FP32 when a_is_f32), zero pad
// rows; diags[j] = squared norm of the stored column, summed in FP64.
hipError_t rbrp_launch_copy_colnorm32(hipStream_t st, const void* A, int a_is_f32, int64_t lda,
                                      int64_t m, int64_t mp, int64_t n, float* Aw, double* diags);
// FP64 working copy (and norms) of an FP32 input.
hipError_t rbrp_launch_copy_colnorm_from32(hipStream_t st, const float* A, int64_t lda, int64_t m,
                                           int64_t mp, int64_t n, double* Aw, double* diags);
// V(:, p) = fp64(Aw(:, J[p])), p < bs->s  /  Aw(:, J[p]) = fp32(V(:, p)) for the
// positions p < bs->s not in kept[0:bs->bt] (all of them when kept is null).
hipError_t rbrp_launch_gather_v64(hipStream_t st, int64_t rows, int b, const RbrpBlockState* bs,
                                  const float* Aw, int64_t ld, const int* J, double* V, int64_t ldv);
hipError_t rbrp_launch_scatter_v32(hipStream_t st, int64_t rows, int b, const RbrpBlockState* bs,
                                   const double* V, int64_t ldv, const int* J, float* Aw, int64_t ld,
                                   const int* kept = nullptr);
// Split-K TN product of FP32 operands with FP64 partials (see k_gemm_tn; a
// single 128-row tile when M > 64);
// A's and B's first columns are acoff + *a_dyn and bcoff + *b_dyn (dyn
// optional), common leading dimension ld (multiple of 4); K and kchunk are
// multiples of 32.
hipError_t rbrp_launch_gemm_tn32(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                                 int64_t K, int nsplit, int64_t kchunk, const float* A,
                                 int64_t acoff, const int* a_dyn, const float* B, int64_t bcoff,
                                 const int* b_dyn, int64_t ld, double* P, int64_t ldp,
                                 int64_t pstride, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// FP32 trailing update with FP64 norm partials over row groups (at most
// ceil(M / 64)); see rbrp_launch_gemm_trail.  M, lda and ldc must be even.
hipError_t rbrp_launch_gemm_trail32(hipStream_t st, int64_t M, DimSpec N, int64_t Nmax, DimSpec K,
                                    int64_t Kmax, const float* A, int64_t lda, int64_t acoff,
                                    const int* acoff_dyn, const float* Bt, int64_t ldbt, float* C,
                                    int64_t ldc, int64_t ccoff, const int* ccoff_dyn,
                                    double* normpart, int64_t ldnp, int64_t* groups = nullptr,
                                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

// ---- k_qrcp.hip ----------------------------------------------------------
// Blocked column-pivoted Householder QR of W (rows x cols, ld) in place to
// min(rows, cols) steps (dgeqp3 / dlaqps organisation: panels of 32 columns,
// partial column norms downdated and recomputed when the downdate is
// rejected): R on and above the diagonal, the reflectors' vectors below it
// with tau, the pivot order in jpvt (device, cols entries).  work: device
// scratch of rbrp_qrcp_work_bytes(rows, cols); host_kb: one pinned int (a
// tall matrix's panel width is read back once per panel).  *panels
// (optional): the panel count.
size_t rbrp_qrcp_work_bytes(int64_t rows, int64_t cols);
hipError_t rbrp_qrcp(hipStream_t st, int64_t rows, int64_t cols, double* W, int64_t ld, int* jpvt,
                     double* tau, void* work, int* host_kb, int* panels = nullptr);

// ---- k_sketch.hip ---------------------------------------------------------
hipError_t rbrp_launch_gaussian(hipStream_t st, uint64_t seed, uint64_t offset, int64_t rows,
                                int64_t cols, double scale, double* out, int64_t ldo);
hipError_t rbrp_launch_gather_cols(hipStream_t st, int64_t rows, int64_t cols, const double* src,
                                   int64_t lds, const int* perm, double* dst, int64_t ldd);
hipError_t rbrp_launch_transpose(hipStream_t st, int64_t rows, int64_t cols, const double* src,
                                 int64_t lds, double* dst, int64_t ldd);
// out(r, j) = Y(j, P[p0 + r]), r < nr, j < nj (transposed gather of sketch columns).
hipError_t rbrp_launch_gather_cols_t(hipStream_t st, int64_t nr, int64_t nj, const double* Y,
                                     int64_t ldy, const int* P, int64_t p0, double* out, int64_t ldo);
// Row interchanges base + j <-> base + ipiv[j] (j < nb, in order) on columns
// 0 .. cols-1 of M and on P (optional).
hipError_t rbrp_launch_swap_rows(hipStream_t st, int nb, int64_t base, const int* ipiv, double* M,
                                 int64_t ldm, int64_t cols, int* P);
hipError_t rbrp_launch_unit_upper(hipStream_t st, int64_t k, double* R, int64_t ld);
// One unpivoted Householder step j of a panel ending at column cend (the
// reflector and its application in one launch; column j is completed by the
// next step or by rbrp_launch_hh_finish); scr holds 3 doubles between steps.
hipError_t rbrp_launch_hh_step(hipStream_t st, int64_t rows, int64_t cend, int64_t j, double* W,
                               int64_t ld, double* tau, double* scr, int has_prev);
hipError_t rbrp_launch_hh_finish(hipStream_t st, int64_t rows, int64_t j, double* W, int64_t ld,
                                 const double* scr);
// Block-reflector pieces of the blocked Householder QR: the panel's explicit
// vectors (unit diagonal, zero above and in the pad rows) and T^T of
// I - V T V^T from S = V^T V and tau (jb <= 64).
hipError_t rbrp_launch_panel_vectors(hipStream_t st, int64_t rows, int64_t rows_pad, int jb,
                                     const double* src, int64_t lds, double* V, int64_t ldv);
hipError_t rbrp_launch_wy_factor(hipStream_t st, int jb, const double* S, int64_t lds, const double* tau,
                                 double* Tt, int64_t ldt);
// vn[c] = ||W(0:rows, c)||^2, one workgroup per column.
hipError_t rbrp_launch_colnorm2(hipStream_t st, int64_t rows, int64_t cols, const double* W,
                                int64_t ld, double* vn);
// LU step j: pivot search down column j, the row swap over columns [c0, ncols),
// the column's scaling.
hipError_t rbrp_launch_lu_pivot(hipStream_t st, int64_t nrows, int64_t c0, int64_t ncols, int64_t j,
                                double* Y, int64_t ld, int* ipiv);
// A panel's interchanges ipiv[j0:j1) applied to the columns outside [j0, j1).
hipError_t rbrp_launch_lu_laswp(hipStream_t st, int64_t ncols, int64_t j0, int64_t j1, double* Y,
                                int64_t ld, const int* ipiv);
// Linv (ld ldi) <- the inverse of the jb x jb unit lower-triangular block at L (jb <= 64).
hipError_t rbrp_launch_unit_lower_inv(hipStream_t st, int jb, const double* L, int64_t ld, double* Linv,
                                      int64_t ldi);
hipError_t rbrp_launch_lu_update(hipStream_t st, int64_t nrows, int64_t ncols, int64_t j, double* Y,
                                 int64_t ld);
