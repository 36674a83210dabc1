// Kernels of the mixed-precision RBRP loop (gfx950): the residual of the
// trailing columns is stored in FP32 and projected / updated with
// v_mfma_f32_16x16x4f32 (FP32 products, FP32 accumulation), while the
// re-orthogonalisation, the pivoted Cholesky / CholQR2 of the sampled block
// and the triangular solve run in FP64 on an FP64 copy of the block.
//
// Operand layout of v_mfma_f32_16x16x4f32 (A 16x4, B 4x16, D 16x16):
//   A[i][k] in lane i + 16k,  B[k][j] in lane j + 16k,
//   D[i][j] in lane j + 16*(i/4), accumulator element i % 4.
#include <algorithm>

#include "rbrp_kernels.h"

typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Norm pass into the FP32 working matrix: Aw(:, j) = fp32([A(0:m, j); 0]),
// diags[j] = sum of the squares of the stored FP32 values, accumulated in FP64
// (so the weights describe exactly the residual that is stored).
// ---------------------------------------------------------------------------
template <typename TIn>
__global__ __launch_bounds__(256) void k_copy_colnorm32(const TIn* __restrict__ A, int64_t lda,
                                                        int64_t m, int64_t mp,
                                                        float* __restrict__ Aw,
                                                        double* __restrict__ diags) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  const TIn* a = A + j * lda;
  float* w = Aw + j * mp;
  double acc0 = 0.0, acc1 = 0.0;
  int64_t r = threadIdx.x;
  for (; r + 256 < m; r += 512) {
    const float v0 = (float)a[r], v1 = (float)a[r + 256];
    w[r] = v0;
    w[r + 256] = v1;
    acc0 = fma((double)v0, (double)v0, acc0);
    acc1 = fma((double)v1, (double)v1, acc1);
  }
  for (; r < mp; r += 256) {
    const float v = r < m ? (float)a[r] : 0.0f;
    w[r] = v;
    acc0 = fma((double)v, (double)v, acc0);
  }
  const double t = rbrp_block_sum(acc0 + acc1, red);
  if (threadIdx.x == 0) diags[j] = t;
}

hipError_t rbrp_launch_copy_colnorm32(hipStream_t st, const void* A, int a_is_f32, int64_t lda,
                                      int64_t m, int64_t mp, int64_t n, float* Aw, double* diags) {
  if (n <= 0) return hipSuccess;
  if (a_is_f32)
    hipLaunchKernelGGL(k_copy_colnorm32<float>, dim3((unsigned)n), dim3(256), 0, st,
                       (const float*)A, lda, m, mp, Aw, diags);
  else
    hipLaunchKernelGGL(k_copy_colnorm32<double>, dim3((unsigned)n), dim3(256), 0, st,
                       (const double*)A, lda, m, mp, Aw, diags);
  return hipGetLastError();
}

// FP64 working copy of an FP32 input (the FP64 loop on FP32-stored data).
__global__ __launch_bounds__(256) void k_copy_colnorm_from32(const float* __restrict__ A,
                                                             int64_t lda, int64_t m, int64_t mp,
                                                             double* __restrict__ Aw,
                                                             double* __restrict__ diags) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  const float* a = A + j * lda;
  double* w = Aw + j * mp;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < mp; r += 256) {
    const double v = r < m ? (double)a[r] : 0.0;
    w[r] = v;
    acc = fma(v, v, acc);
  }
  const double t = rbrp_block_sum(acc, red);
  if (threadIdx.x == 0) diags[j] = t;
}

hipError_t rbrp_launch_copy_colnorm_from32(hipStream_t st, const float* A, int64_t lda, int64_t m,
                                           int64_t mp, int64_t n, double* Aw, double* diags) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_colnorm_from32, dim3((unsigned)n), dim3(256), 0, st, A, lda, m, mp, Aw,
                     diags);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sampled block <-> FP64 copy.  V(:, p) = fp64(Aw(:, J[p])) for p < s, and the
// re-orthogonalised columns written back, Aw(:, J[p]) = fp32(V(:, p)).
// ---------------------------------------------------------------------------
__global__ void k_gather_v64(int64_t rows, const RbrpBlockState* bs, const float* __restrict__ Aw,
                             int64_t ld, const int* __restrict__ J, double* __restrict__ V,
                             int64_t ldv) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int p = blockIdx.y;
  if (r >= rows || p >= bs->s) return;
  V[r + (int64_t)p * ldv] = (double)Aw[r + (int64_t)J[p] * ld];
}

// (kept: the sampled positions of the kept columns, kept[0:bs->bt]; those
// columns become the block's basis, so only the filtered-out ones are written.
// A workgroup strides over the rows, so the membership test runs once per
// column and workgroup, not per 256 rows; 256 threads.)
__global__ void k_scatter_v32(int64_t rows, const RbrpBlockState* bs, const double* __restrict__ V,
                              int64_t ldv, const int* __restrict__ J, const int* __restrict__ kept,
                              float* __restrict__ Aw, int64_t ld) {
  const int p = blockIdx.y;
  if (p >= bs->s) return;  // uniform per workgroup
  // one list entry per thread (bt <= 128 < 256), a block-wide OR
  const int hit = kept && (int)threadIdx.x < bs->bt && kept[threadIdx.x] == p;
  if (__syncthreads_or(hit)) return;  // uniform per workgroup
  const double* v = V + (int64_t)p * ldv;
  float* a = Aw + (int64_t)J[p] * ld;
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < rows; r += (int64_t)gridDim.x * 256)
    a[r] = (float)v[r];
}

hipError_t rbrp_launch_gather_v64(hipStream_t st, int64_t rows, int b, const RbrpBlockState* bs,
                                  const float* Aw, int64_t ld, const int* J, double* V,
                                  int64_t ldv) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_v64, dim3((unsigned)((rows + 255) / 256), (unsigned)b), dim3(256), 0,
                     st, rows, bs, Aw, ld, J, V, ldv);
  return hipGetLastError();
}

hipError_t rbrp_launch_scatter_v32(hipStream_t st, int64_t rows, int b, const RbrpBlockState* bs,
                                   const double* V, int64_t ldv, const int* J, float* Aw,
                                   int64_t ld, const int* kept) {
  if (rows <= 0) return hipSuccess;
  const int64_t gx = std::min<int64_t>((rows + 255) / 256, 256);
  hipLaunchKernelGGL(k_scatter_v32, dim3((unsigned)gx, (unsigned)b), dim3(256), 0, st, rows, bs, V,
                     ldv, J, kept, Aw, ld);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split-K TN in FP32: P[z](r, c) = sum_{kk in chunk z} A(kk, r) B(kk, c) with
// FP32 MFMA products and accumulation, stored as FP64 partials (the split sum
// then runs in FP64).  A (the orthonormal block) and B (the trailing matrix)
// are column-major FP32 with a common leading dimension; their first columns
// are acoff + *a_dyn and bcoff + *b_dyn.
//
// Workgroup tile: MT = 64 or 128 rows (128 when A is wider than 64 columns, so
// that B is streamed once per split) x 128 columns (an XCD-aware task); wave w
// owns columns 32w .. 32w+31 (two 16-column MFMA blocks) against the MT / 16
// row blocks of A.  Each 32-row step of both operands is staged through
// double-buffered LDS ([column][row], one 128-byte run per column and step,
// its eight 16-byte units XOR-swizzled with the column index mod 8) with two
// steps of global loads in flight: the loads of step k+2 are issued before
// the MFMAs of step k and step k+1's registers are written to LDS after them
// (one barrier per step).  Lane (j, g) reads rows 8g .. 8g+7 of its column as
// two ds_read_b128, and MFMA q of the step consumes element q: the eight
// MFMAs cover the 32 rows in the order 8g + q.  The MFMA computes the
// transposed block (B's columns as its rows), so lane l holds
// P(r0 + l%16, c0 + 4(l/16) + v) and each partial store instruction writes
// four 128-byte runs of P.  kchunk and K are multiples of 32; rows >= the
// true height are zero.  Every load and LDS store is unconditional (a step
// past the chunk re-reads its last step and writes the buffer not read), so
// the vector-memory waits count only the loads issued before each store.
// ---------------------------------------------------------------------------
#define TN32_KS 32
// LDS offset (floats) of 16-byte unit u (rows 4u .. 4u+3) of staged column col
#define TN32_AT(col, u) ((col) * TN32_KS + 4 * ((u) ^ ((col) & 7)))

template <int MT>
__global__ __launch_bounds__(256, 2) void k_gemm_tn32(DimSpec dM, DimSpec dN, int64_t K,
                                                      int64_t kchunk, const float* __restrict__ A0,
                                                      int64_t acoff, const int* a_dyn,
                                                      const float* __restrict__ B0, int64_t bcoff,
                                                      const int* b_dyn, int64_t ld,
                                                      double* __restrict__ P, int64_t ldp,
                                                      int64_t pstride, int gx, int gy, int gz) {
  constexpr int MB = MT / 16;  // 16-row MFMA blocks of A per tile
  // task = N tile + gx * (M tile + gy * split), XCD-aware (the tiles of one
  // split share its rows of A through one L2)
  const int64_t task = rbrp_xcd_task(blockIdx.x, (int64_t)gx * gy * gz);
  if (task >= (int64_t)gx * gy * gz) return;
  const int64_t bx = task % gx, by = (task / gx) % gy, bz = task / ((int64_t)gx * gy);
  const int64_t M = dim_eval(dM), N = dim_eval(dN);
  const int64_t n0 = bx * 128, m0 = by * MT;
  if (m0 >= M || n0 >= N) return;  // uniform per workgroup
  const int64_t kb0 = bz * kchunk;
  const int64_t kb1 = min(K, kb0 + kchunk);
  if (kb0 >= kb1) return;  // empty chunk (uniform); its partials are never read

  __shared__ float As[2][MT * TN32_KS];   // [buf][col][row], swizzled
  __shared__ float Bs[2][128 * TN32_KS];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lj = lane & 15, lg = lane >> 4;
  const float* A = A0 + (acoff + (a_dyn ? *a_dyn : 0)) * ld;
  const float* B = B0 + (bcoff + (b_dyn ? *b_dyn : 0)) * ld;
  // Stagers.  Out-of-range columns are clamped to the last valid one: they
  // only feed output rows / columns that are never stored.
  constexpr int AQ = MT / 8;  // A rows per thread and step: 8 (MT = 64) or 16
  const int ac = tid / (256 / MT), aq = (tid % (256 / MT)) * AQ;  // A: column ac, rows aq ..
  const int bc = tid >> 1, bq = (tid & 1) * 16;   // B: column bc, rows bq .. bq+15
  const float* ap = A + min(m0 + ac, M - 1) * ld + aq;
  const float* bp = B + min(n0 + bc, N - 1) * ld + bq;
  const int au = aq >> 2, bu = bq >> 2;  // first 16-byte unit of the stager
  // row blocks holding rows < M (A's width is often below the tile height)
  const int mba = (int)min((int64_t)MB, (M - m0 + 15) / 16);

  v4f acc[MB][2];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  // two sets of staging registers (named scalars: small arrays here get
  // placed in LDS); the A units past AQ / 4 exist only for MT = 128
#define TN32_LOAD(S, kk)                                                  \
  do {                                                                    \
    S##a0 = *(const float4*)(ap + (kk));                                  \
    S##a1 = *(const float4*)(ap + (kk) + 4);                              \
    if (MT == 128) {                                                      \
      S##a2 = *(const float4*)(ap + (kk) + 8);                            \
      S##a3 = *(const float4*)(ap + (kk) + 12);                           \
    }                                                                     \
    S##b0 = *(const float4*)(bp + (kk));                                  \
    S##b1 = *(const float4*)(bp + (kk) + 4);                              \
    S##b2 = *(const float4*)(bp + (kk) + 8);                              \
    S##b3 = *(const float4*)(bp + (kk) + 12);                             \
  } while (0)
#define TN32_STORE(S, bb)                                                 \
  do {                                                                    \
    *(float4*)&As[bb][TN32_AT(ac, au)] = S##a0;                           \
    *(float4*)&As[bb][TN32_AT(ac, au + 1)] = S##a1;                       \
    if (MT == 128) {                                                      \
      *(float4*)&As[bb][TN32_AT(ac, au + 2)] = S##a2;                     \
      *(float4*)&As[bb][TN32_AT(ac, au + 3)] = S##a3;                     \
    }                                                                     \
    *(float4*)&Bs[bb][TN32_AT(bc, bu)] = S##b0;                           \
    *(float4*)&Bs[bb][TN32_AT(bc, bu + 1)] = S##b1;                       \
    *(float4*)&Bs[bb][TN32_AT(bc, bu + 2)] = S##b2;                       \
    *(float4*)&Bs[bb][TN32_AT(bc, bu + 3)] = S##b3;                       \
  } while (0)
#define TN32_STEP(bb)                                                                         \
  do {                                                                                        \
    const float* as = &As[bb][0];                                                             \
    const float* bs = &Bs[bb][0];                                                             \
    float bv[2][8];                                                                           \
    _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                           \
      const float4 x0 = *(const float4*)&bs[TN32_AT(32 * wid + 16 * h + lj, 2 * lg)];         \
      const float4 x1 = *(const float4*)&bs[TN32_AT(32 * wid + 16 * h + lj, 2 * lg + 1)];     \
      bv[h][0] = x0.x; bv[h][1] = x0.y; bv[h][2] = x0.z; bv[h][3] = x0.w;                     \
      bv[h][4] = x1.x; bv[h][5] = x1.y; bv[h][6] = x1.z; bv[h][7] = x1.w;                     \
    }                                                                                         \
    _Pragma("unroll") for (int mb = 0; mb < MB; ++mb) {                                       \
      if (mb >= mba) break; /* uniform */                                                     \
      const float4 a0 = *(const float4*)&as[TN32_AT(16 * mb + lj, 2 * lg)];                   \
      const float4 a1 = *(const float4*)&as[TN32_AT(16 * mb + lj, 2 * lg + 1)];               \
      const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};                   \
      _Pragma("unroll") for (int q = 0; q < 8; ++q) {                                         \
        acc[mb][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[0][q], av[q], acc[mb][0], 0, 0, 0); \
        acc[mb][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[1][q], av[q], acc[mb][1], 0, 0, 0); \
      }                                                                                       \
    }                                                                                         \
  } while (0)

  const int64_t klast = kb1 - TN32_KS;
  float4 Xa0, Xa1, Xa2, Xa3, Xb0, Xb1, Xb2, Xb3;
  float4 Ya0, Ya1, Ya2, Ya3, Yb0, Yb1, Yb2, Yb3;
  TN32_LOAD(X, kb0);
  TN32_STORE(X, 0);
  TN32_LOAD(X, min(kb0 + TN32_KS, klast));
  __syncthreads();
  // whole pairs of steps, then an odd last step: no branch inside the loop
  // body.  Step k reads buffer k & 1; X holds step k+1 at even k, Y at odd k.
  const int64_t nsteps = (kb1 - kb0) / TN32_KS;
  int64_t kb = kb0;
  for (int64_t s2 = 2; s2 <= nsteps; s2 += 2, kb += 2 * TN32_KS) {
    TN32_LOAD(Y, min(kb + 2 * TN32_KS, klast));
    __builtin_amdgcn_sched_barrier(0);
    TN32_STEP(0);
    __builtin_amdgcn_sched_barrier(0);
    TN32_STORE(X, 1);
    __syncthreads();
    TN32_LOAD(X, min(kb + 3 * TN32_KS, klast));
    __builtin_amdgcn_sched_barrier(0);
    TN32_STEP(1);
    __builtin_amdgcn_sched_barrier(0);
    TN32_STORE(Y, 0);
    __syncthreads();
  }
  if (nsteps & 1) TN32_STEP(0);
#undef TN32_LOAD
#undef TN32_STORE
#undef TN32_STEP
  double* Pz = P + bz * pstride;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = m0 + 16 * mb + lj;
        const int64_t c = n0 + 32 * wid + 16 * h + 4 * lg + v;
        if (r < M && c < N) Pz[r + c * ldp] = (double)acc[mb][h][v];
      }
}

int rbrp_gemm_tn32_resident(int wide) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per, wide ? k_gemm_tn32<128> : k_gemm_tn32<64>, 256, 0) != hipSuccess)
    return 0;
  return cus * per;
}

hipError_t rbrp_launch_gemm_tn32(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                                 int64_t K, int nsplit, int64_t kchunk, const float* A,
                                 int64_t acoff, const int* a_dyn, const float* B, int64_t bcoff,
                                 const int* b_dyn, int64_t ld, double* P, int64_t ldp,
                                 int64_t pstride, hipEvent_t e0, hipEvent_t e1) {
  if (Mmax <= 0 || Nmax <= 0 || nsplit <= 0) return hipSuccess;
  if (kchunk % TN32_KS != 0 || K % TN32_KS != 0 || ld % 4 != 0) return hipErrorInvalidValue;
  // the whole A block in one 128-row tile when it is wider than 64 (the
  // trailing columns B are then streamed once per split, not once per tile)
  const bool wide = Mmax > 64;
  const int64_t gx = (Nmax + 127) / 128, gy = wide ? (Mmax + 127) / 128 : (Mmax + 63) / 64;
  dim3 grid((unsigned)rbrp_xcd_grid(gx * gy * nsplit));
  if (wide)
    RBRP_LAUNCH_EV(e0, e1, k_gemm_tn32<128>, grid, dim3(256), 0, st, M, N, K, kchunk, A, acoff,
                   a_dyn, B, bcoff, b_dyn, ld, P, ldp, pstride, (int)gx, (int)gy, nsplit);
  else
    RBRP_LAUNCH_EV(e0, e1, k_gemm_tn32<64>, grid, dim3(256), 0, st, M, N, K, kchunk, A, acoff,
                   a_dyn, B, bcoff, b_dyn, ld, P, ldp, pstride, (int)gx, (int)gy, nsplit);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP32 trailing update with fused norms: C(r, c) -= sum_k A(r, k) Bt[k * ldbt + c]
// for r < M, c < N, K <= Kmax <= 128; optional squared column norms of the new
// C, one FP64 partial per workgroup row group (FP32 squares accumulated per
// lane over the sweep, then summed in FP64 across lanes and waves).
//
// The MFMA computes D = B^T A^T (D[i][j] = C(row j, column i) of a 16x16
// block) so that C needs no LDS staging.  The 16 "j" rows of the block are
// chosen as r0 + 2l + jb (l = 0..15, jb = 0..1) over the two row blocks a
// wave owns, so that lane l holds rows r0 + 2l and r0 + 2l + 1 of the same
// four columns: every C access is a float2, and one instruction moves four
// 128-byte column runs.  The A fragments follow the same row order (one
// float2 per k per lane).
//
// Workgroup = 4 waves stacked along rows (32 rows each), 128 rows x 32
// columns per tile, sweeping T consecutive row tiles; B (K x 32) in LDS; the
// next tile's C is prefetched after the current tile's A fragments.
// ---------------------------------------------------------------------------
template <int KSTEPS, bool NORMS>
__global__ __launch_bounds__(256) void k_gemm_trail32(int64_t M, DimSpec dN, DimSpec dK, int T,
                                                      const float* __restrict__ A0, int64_t lda,
                                                      int64_t acoff, const int* acoff_dyn,
                                                      const float* __restrict__ Bt, int64_t ldbt,
                                                      float* __restrict__ C0, int64_t ldc,
                                                      int64_t ccoff, const int* ccoff_dyn,
                                                      double* __restrict__ normpart, int64_t ldnp) {
  __shared__ float Ws[4 * KSTEPS * 32];  // [k][32]
  __shared__ double red[4][32];
  const int64_t N = dim_eval(dN), K = dim_eval(dK);
  const int64_t c0 = (int64_t)blockIdx.x * 32;
  const int64_t rbase = (int64_t)blockIdx.y * T * 128;
  if (c0 >= N || rbase >= M || K <= 0) return;  // uniform per workgroup
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const float* A = A0 + (acoff + (acoff_dyn ? *acoff_dyn : 0)) * lda;
  float* C = C0 + (ccoff + (ccoff_dyn ? *ccoff_dyn : 0) + c0) * ldc;
  for (int idx = threadIdx.x; idx < 4 * KSTEPS * 32; idx += 256) {
    const int k = idx >> 5, c = idx & 31;
    Ws[idx] = (k < K && c0 + c < N) ? Bt[(int64_t)k * ldbt + c0 + c] : 0.0f;
  }
  __syncthreads();
  const int ncol = (int)min((int64_t)32, N - c0);
  // Loads are unconditional, with the column clamped to the last valid one
  // and the row to M - 2 (M is even, rows are read in pairs): the clamped
  // values only reach results that are not stored, and A's clamped columns
  // (k >= K) meet zero rows of Ws.
  const float* ccol[2][4];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int v = 0; v < 4; ++v) ccol[ib][v] = C + (int64_t)min(16 * ib + 4 * lk + v, ncol - 1) * ldc;
  const float* acol[KSTEPS];
#pragma unroll
  for (int q = 0; q < KSTEPS; ++q) acol[q] = A + min((int64_t)(4 * q + lk), K - 1) * lda;

#define TRAIL32_LOADC(R0, X00, X01, X02, X03, X10, X11, X12, X13)     \
  {                                                                    \
    const int64_t rr = min((R0) + 32 * wid + 2 * li, M - 2);          \
    X00 = *(const float2*)(ccol[0][0] + rr);                           \
    X01 = *(const float2*)(ccol[0][1] + rr);                           \
    X02 = *(const float2*)(ccol[0][2] + rr);                           \
    X03 = *(const float2*)(ccol[0][3] + rr);                           \
    X10 = *(const float2*)(ccol[1][0] + rr);                           \
    X11 = *(const float2*)(ccol[1][1] + rr);                           \
    X12 = *(const float2*)(ccol[1][2] + rr);                           \
    X13 = *(const float2*)(ccol[1][3] + rr);                           \
  }
  // squared norms of the new C over the whole sweep, reduced once at its end
  float nsq[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
  float2 a00, a01, a02, a03, a10, a11, a12, a13;  // C of the current tile
  float2 n00, n01, n02, n03, n10, n11, n12, n13;  // C of the next tile
  TRAIL32_LOADC(rbase, a00, a01, a02, a03, a10, a11, a12, a13);
  for (int t = 0; t < T; ++t) {
    const int64_t r0 = rbase + (int64_t)t * 128;
    if (r0 >= M) break;  // uniform per workgroup
    const int64_t r = r0 + 32 * wid + 2 * li;
    const int64_t rc = min(r, M - 2);
    float2 fa[KSTEPS];
#pragma unroll
    for (int q = 0; q < KSTEPS; ++q) fa[q] = *(const float2*)(acol[q] + rc);
    if (t + 1 < T) TRAIL32_LOADC(r0 + 128, n00, n01, n02, n03, n10, n11, n12, n13);
    v4f acc00 = {0.f, 0.f, 0.f, 0.f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;  // [jb][ib]
#pragma unroll
    for (int q = 0; q < KSTEPS; ++q) {
      const int k = 4 * q + lk;
      const float b0 = Ws[k * 32 + li], b1 = Ws[k * 32 + 16 + li];
      acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0, fa[q].x, acc00, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(b1, fa[q].x, acc01, 0, 0, 0);
      acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0, fa[q].y, acc10, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(b1, fa[q].y, acc11, 0, 0, 0);
    }
    // lane (li, lk) holds D[4lk + v][li]: column 16ib + 4lk + v, rows r (jb 0) and r + 1 (jb 1)
    const bool rok = r < M;
#define TRAIL32_STORE(IB, V, X)                                                 \
  {                                                                             \
    const int cc = 16 * (IB) + 4 * lk + (V);                                    \
    float2 nv;                                                                  \
    nv.x = X.x - acc0##IB[V];                                                   \
    nv.y = X.y - acc1##IB[V];                                                   \
    if (cc < ncol && rok) {                                                     \
      *(float2*)(C + r + (int64_t)cc * ldc) = nv;                               \
      nsq[IB][V] = fmaf(nv.x, nv.x, fmaf(nv.y, nv.y, nsq[IB][V]));              \
    }                                                                           \
  }
    TRAIL32_STORE(0, 0, a00);
    TRAIL32_STORE(0, 1, a01);
    TRAIL32_STORE(0, 2, a02);
    TRAIL32_STORE(0, 3, a03);
    TRAIL32_STORE(1, 0, a10);
    TRAIL32_STORE(1, 1, a11);
    TRAIL32_STORE(1, 2, a12);
    TRAIL32_STORE(1, 3, a13);
#undef TRAIL32_STORE
    a00 = n00; a01 = n01; a02 = n02; a03 = n03;
    a10 = n10; a11 = n11; a12 = n12; a13 = n13;
  }
#undef TRAIL32_LOADC
  if (NORMS) {
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double x = (double)nsq[ib][v];  // FP64 across lanes and waves
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        x += __shfl_xor(x, 8, 64);
        if (li == 0) red[wid][16 * ib + 4 * lk + v] = x;
      }
    __syncthreads();
    if (wid == 0 && lane < ncol)
      normpart[(int64_t)blockIdx.y * ldnp + c0 + lane] =
          (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
  }
}

// ---------------------------------------------------------------------------
// FP32 trailing update for K <= KB (64 or 128) and M a multiple of 64 (the
// mixed loop's shape), the FP32 form of k_gemm_upd64: 128 columns per
// workgroup (4 waves side by side, 32 columns each), row tiles of 64 rows
// whose 64 x KB block of A is staged once in double-buffered LDS and shared by
// the four waves, each wave's KB x 32 block of B in registers for the whole
// sweep.  A wave takes a row tile as two 32-row sub-blocks (lane l owns rows
// 2(l%16) and 2(l%16)+1 of a sub-block, as float2) and prefetches the next
// sub-block's C.  LDS holds A as row pairs ([pair][k][2], 2 KB + 4 floats per
// pair) so that a lane reads its KB / 4 k values of both rows with 16-byte
// loads; MFMA q uses k = (KB / 4) (lane / 16) + q.  Norm partials: FP32 squares
// per lane over the sweep, summed in FP64 across the 16 lanes of a column.
// ---------------------------------------------------------------------------
template <int KB, bool NORMS>
__global__ __launch_bounds__(256, 2) void k_gemm_upd32(int64_t M, DimSpec dN, DimSpec dK, int T,
                                                       int ncol, int nrg,
                                                       const float* __restrict__ A0, int64_t lda,
                                                       int64_t acoff, const int* acoff_dyn,
                                                       const float* __restrict__ Bt, int64_t ldbt,
                                                       float* __restrict__ C0, int64_t ldc,
                                                       int64_t ccoff, const int* ccoff_dyn,
                                                       double* __restrict__ normpart, int64_t ldnp) {
  constexpr int KQ = KB / 4;          // k values per lane (and MFMAs per column block)
  constexpr int LDP = 2 * KB + 4;     // floats per row pair: 16 lanes on distinct banks
  constexpr int TPC = 256 / KB;       // stager threads per column of A
  constexpr int RPT = 64 / TPC;       // rows per stager thread
  __shared__ float As[2][32 * LDP];
  const int64_t N = dim_eval(dN), K = dim_eval(dK);
  // task = column tile + ncol * row group, XCD-aware (a row group's column
  // tiles share its A rows through one L2)
  const int64_t task = rbrp_xcd_task(blockIdx.x, (int64_t)ncol * nrg);
  if (task >= (int64_t)ncol * nrg) return;
  const int64_t bx = task % ncol, by = task / ncol;
  const int64_t c0 = bx * 128;
  const int64_t rbase = by * T * 64;
  if (c0 >= N || rbase >= M || K <= 0) return;  // uniform per workgroup
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const float* A = A0 + (acoff + (acoff_dyn ? *acoff_dyn : 0)) * lda;
  const int64_t cw = c0 + 32 * wid;
  const int ncw = (int)max((int64_t)0, min((int64_t)32, N - cw));
  float* C = C0 + (ccoff + (ccoff_dyn ? *ccoff_dyn : 0) + (ncw > 0 ? cw : c0)) * ldc;
  const int cmax = max(ncw, 1) - 1;
  // MFMA q of a sub-block takes k = 4q .. 4q+3 (k = 4q + lk in lane group lk),
  // so every q with 4q >= K is skipped (the kept count is often below KB)
  const int qa = (int)min((int64_t)KQ, ((K + 7) / 8) * 2);  // MFMA steps in use, even
  float b0[KQ], b1[KQ];  // B(4q + lk, cw + li) and B(4q + lk, cw + 16 + li)
#pragma unroll
  for (int q = 0; q < KQ; ++q) {
    const int64_t k = 4 * q + lk;
    b0[q] = (k < K && li < ncw) ? Bt[k * ldbt + cw + li] : 0.0f;
    b1[q] = (k < K && 16 + li < ncw) ? Bt[k * ldbt + cw + 16 + li] : 0.0f;
  }
  // A stager: column ka (clamped to K - 1: B is zero there), rows ar .. ar + RPT - 1 of a tile
  const int ka = tid / TPC, ar = RPT * (tid % TPC);
  const float* acol = A + min((int64_t)ka, K - 1) * lda + ar;
  float4 sv[RPT / 4];
#define UPD32_LOADA(row0)                                                             \
  do {                                                                                \
    const float* ap_ = acol + (row0);                                                 \
    _Pragma("unroll") for (int j = 0; j < RPT / 4; ++j) sv[j] = *(const float4*)(ap_ + 4 * j); \
  } while (0)
  // column ka sits at k position (ka % 4) KQ + ka / 4 of its row pair, so the k
  // values 4q + lk of lane group lk are contiguous
#define UPD32_STOREA(bb)                                                              \
  do {                                                                                \
    float* d_ = &As[bb][0] + (ar / 2) * LDP + 2 * ((ka & 3) * KQ + (ka >> 2));        \
    _Pragma("unroll") for (int j = 0; j < RPT / 4; ++j) {                             \
      *(float2*)(d_ + (2 * j) * LDP) = make_float2(sv[j].x, sv[j].y);                 \
      *(float2*)(d_ + (2 * j + 1) * LDP) = make_float2(sv[j].z, sv[j].w);             \
    }                                                                                 \
  } while (0)
  // C of one 32-row sub-block: cv[h][v] = C(r + 2 li .. +1, 16 h + 4 lk + v)
#define UPD32_LOADC(cv, r)                                                            \
  do {                                                                                \
    _Pragma("unroll") for (int h = 0; h < 2; ++h)                                     \
      _Pragma("unroll") for (int v = 0; v < 4; ++v)                                   \
        cv[h][v] = *(const float2*)(C + (r) + 2 * li +                                \
                                    (int64_t)min(16 * h + 4 * lk + v, cmax) * ldc);   \
  } while (0)

  float nsq[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int v = 0; v < 4; ++v) nsq[h][v] = 0.0f;
  UPD32_LOADA(rbase);
  UPD32_STOREA(0);
  float2 cc[2][4], cn[2][4];
  UPD32_LOADC(cc, rbase);
  __syncthreads();
  int buf = 0;
  for (int t = 0; t < T; ++t) {
    const int64_t row0 = rbase + (int64_t)t * 64;
    if (row0 >= M) break;  // uniform
    const bool more = t + 1 < T && row0 + 64 < M;
    const int64_t rnext = min(row0 + 64, M - 64);
    UPD32_LOADA(rnext);
    const float* as = &As[buf][0];
#pragma unroll
    for (int sb = 0; sb < 2; ++sb) {
      const int64_t r = row0 + 32 * sb;
      UPD32_LOADC(cn, sb == 0 ? r + 32 : rnext);
      const float* apair = as + (16 * sb + li) * LDP + 2 * KQ * lk;  // k = 4q + lk, two rows
      v4f a00 = (v4f){0.f, 0.f, 0.f, 0.f}, a01 = a00, a10 = a00, a11 = a00;  // [row of pair][h]
#pragma unroll
      for (int q = 0; q < KQ; q += 2) {
        if (q >= qa) break;  // uniform
        // (4q + lk, row 0), (4q + lk, row 1), (4q + 4 + lk, row 0), (4q + 4 + lk, row 1)
        const float4 x = *(const float4*)(apair + 2 * q);
        a00 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[q], x.x, a00, 0, 0, 0);
        a01 = __builtin_amdgcn_mfma_f32_16x16x4f32(b1[q], x.x, a01, 0, 0, 0);
        a10 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[q], x.y, a10, 0, 0, 0);
        a11 = __builtin_amdgcn_mfma_f32_16x16x4f32(b1[q], x.y, a11, 0, 0, 0);
        a00 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[q + 1], x.z, a00, 0, 0, 0);
        a01 = __builtin_amdgcn_mfma_f32_16x16x4f32(b1[q + 1], x.z, a01, 0, 0, 0);
        a10 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[q + 1], x.w, a10, 0, 0, 0);
        a11 = __builtin_amdgcn_mfma_f32_16x16x4f32(b1[q + 1], x.w, a11, 0, 0, 0);
      }
      // lane (li, lk) holds D[4 lk + v][li]: column 16 h + 4 lk + v, rows r + 2 li (+1)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ca = 4 * lk + v, cb = 16 + 4 * lk + v;
        const float2 na = make_float2(cc[0][v].x - a00[v], cc[0][v].y - a10[v]);
        const float2 nb = make_float2(cc[1][v].x - a01[v], cc[1][v].y - a11[v]);
        if (ca < ncw) {
          *(float2*)(C + r + 2 * li + (int64_t)ca * ldc) = na;
          if (NORMS) nsq[0][v] = fmaf(na.x, na.x, fmaf(na.y, na.y, nsq[0][v]));
        }
        if (cb < ncw) {
          *(float2*)(C + r + 2 * li + (int64_t)cb * ldc) = nb;
          if (NORMS) nsq[1][v] = fmaf(nb.x, nb.x, fmaf(nb.y, nb.y, nsq[1][v]));
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < 4; ++v) cc[h][v] = cn[h][v];
    }
    if (more) UPD32_STOREA(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#undef UPD32_LOADA
#undef UPD32_STOREA
#undef UPD32_LOADC
  if (NORMS) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double x = (double)nsq[h][v];  // FP64 across lanes
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        x += __shfl_xor(x, 8, 64);
        const int cl = 16 * h + 4 * lk + v;
        if (li == 0 && cl < ncw) normpart[by * ldnp + cw + cl] = x;
      }
  }
}

hipError_t rbrp_launch_gemm_trail32(hipStream_t st, int64_t M, DimSpec N, int64_t Nmax, DimSpec K,
                                    int64_t Kmax, const float* A, int64_t lda, int64_t acoff,
                                    const int* acoff_dyn, const float* Bt, int64_t ldbt, float* C,
                                    int64_t ldc, int64_t ccoff, const int* ccoff_dyn,
                                    double* normpart, int64_t ldnp, int64_t* groups,
                                    hipEvent_t e0, hipEvent_t e1) {
  if (groups) *groups = 0;
  if (M <= 0 || Nmax <= 0) return hipSuccess;
  if (Kmax > 128 || (M & 1) || (lda & 1) || (ldc & 1)) return hipErrorInvalidValue;
  if (M % 64 == 0 && lda % 4 == 0) {
    // 128-column workgroups with the A block shared through LDS
    const int64_t ncol = (Nmax + 127) / 128, nrow = M / 64;
    const int64_t T = std::max<int64_t>(1, std::min<int64_t>(8, (ncol * nrow) / 512));
    const int64_t nrg = (nrow + T - 1) / T;
    dim3 grid((unsigned)rbrp_xcd_grid(ncol * nrg));
    if (groups) *groups = nrg;
#define UPD32_LAUNCH(KB, NM)                                                                     \
  RBRP_LAUNCH_EV(e0, e1, (k_gemm_upd32<KB, NM>), grid, dim3(256), 0, st, M, N, K, (int)T, (int)ncol, \
                     (int)nrg, A, lda, acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff, ccoff_dyn,     \
                     normpart, ldnp)
    if (Kmax <= 64) {
      if (normpart)
        UPD32_LAUNCH(64, true);
      else
        UPD32_LAUNCH(64, false);
    } else {
      if (normpart)
        UPD32_LAUNCH(128, true);
      else
        UPD32_LAUNCH(128, false);
    }
#undef UPD32_LAUNCH
    return hipGetLastError();
  }
  const bool small = Kmax <= 64;
  const int64_t ncol = (Nmax + 31) / 32, nrow = (M + 127) / 128;
  int64_t T = std::max<int64_t>(1, std::min<int64_t>(8, (ncol * nrow) / 2048));
  dim3 grid((unsigned)ncol, (unsigned)((nrow + T - 1) / T));
  if (groups) *groups = (int64_t)grid.y;
#define TRAIL32_LAUNCH(KS, NM)                                                                   \
  RBRP_LAUNCH_EV(e0, e1, (k_gemm_trail32<KS, NM>), grid, dim3(256), 0, st, M, N, K, (int)T, A, lda,  \
                     acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff, ccoff_dyn, normpart, ldnp)
  if (small) {
    if (normpart)
      TRAIL32_LAUNCH(16, true);
    else
      TRAIL32_LAUNCH(16, false);
  } else {
    if (normpart)
      TRAIL32_LAUNCH(32, true);
    else
      TRAIL32_LAUNCH(32, false);
  }
#undef TRAIL32_LAUNCH
  return hipGetLastError();
}
