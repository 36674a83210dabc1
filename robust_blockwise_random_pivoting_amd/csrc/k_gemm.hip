// FP64 MFMA GEMM kernels for the RBRP loop (gfx950).
//
// Both kernels use v_mfma_f64_16x16x4f64 on 64x64 workgroup tiles: four
// waves in a 2x2 arrangement, each owning a 32x32 sub-tile (2x2 MFMA blocks),
// with 16-deep K tiles staged through LDS and the next K tile prefetched into
// registers while the current one is multiplied.
//
// Operand layout of v_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16):
//   A[i][k] in lane i + 16k,  B[k][j] in lane j + 16k,
//   D[i][j] in lane j + 16*(i%4), accumulator element i / 4
//   (measured on gfx950 with a one-hot probe of the A operand).
#include <algorithm>

#include "rbrp_kernels.h"

#define KT 16

// Trailing-update launch shape: rows per wave / 16 (JB), longest row sweep per
// workgroup (TMAX) and the grid size the sweep length aims at (WGS workgroups).
#ifndef RBRP_TRAIL_JB
#define RBRP_TRAIL_JB 1
#endif
#ifndef RBRP_TRAIL_TMAX
#define RBRP_TRAIL_TMAX 8
#endif
#ifndef RBRP_TRAIL_WGS
#define RBRP_TRAIL_WGS 2048
#endif

__device__ inline int mf_row(int lane, int v) { return 4 * v + (lane >> 4); }
__device__ inline int mf_col(int lane) { return lane & 15; }

// ---------------------------------------------------------------------------
// Split-K TN: P[z](r, c) = sum_{kk in [z*kchunk, (z+1)*kchunk) ∩ [0,K)} A(kk, r) B(kk, c)
// The reduction dimension is the contiguous one of both operands (column
// dot products of two tall matrices).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm_tn(DimSpec dM, DimSpec dN, int64_t K, int64_t kchunk,
                                                 MatRef A, MatRef B, double* __restrict__ P,
                                                 int64_t ldp, int64_t pstride) {
  const int64_t M = dim_eval(dM), N = dim_eval(dN);
  const int64_t m0 = (int64_t)blockIdx.y * 64, n0 = (int64_t)blockIdx.x * 64;
  if (m0 >= M || n0 >= N) return;
  const int64_t kb0 = (int64_t)blockIdx.z * kchunk;
  const int64_t kb1 = min(K, kb0 + kchunk);

  __shared__ double As[64 * (KT + 1)];  // [col][k]
  __shared__ double Bs[64 * (KT + 1)];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int lc = tid >> 2, kq = (tid & 3) * 4;
  const double* ap = (m0 + lc < M) ? mat_col(A, m0 + lc) : nullptr;
  const double* bp = (n0 + lc < N) ? mat_col(B, n0 + lc) : nullptr;

  double ra[4], rb[4];
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

  if (kb0 < kb1) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra[q] = ap ? ap[kb0 + kq + q] : 0.0;
      rb[q] = bp ? bp[kb0 + kq + q] : 0.0;
    }
  }
  for (int64_t kb = kb0; kb < kb1; kb += KT) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[lc * (KT + 1) + kq + q] = ra[q];
      Bs[lc * (KT + 1) + kq + q] = rb[q];
    }
    __syncthreads();
    if (kb + KT < kb1) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ra[q] = ap ? ap[kb + KT + kq + q] : 0.0;
        rb[q] = bp ? bp[kb + KT + kq + q] : 0.0;
      }
    }
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4) {
      const int ko = kk + (lane >> 4);
      const double a0 = As[(wm * 32 + (lane & 15)) * (KT + 1) + ko];
      const double a1 = As[(wm * 32 + 16 + (lane & 15)) * (KT + 1) + ko];
      const double b0 = Bs[(wn * 32 + (lane & 15)) * (KT + 1) + ko];
      const double b1 = Bs[(wn * 32 + 16 + (lane & 15)) * (KT + 1) + ko];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  double* Pz = P + (int64_t)blockIdx.z * pstride;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = m0 + wm * 32 + mi * 16 + mf_row(lane, v);
        const int64_t c = n0 + wn * 32 + ni * 16 + mf_col(lane);
        if (r < M && c < N) Pz[r + c * ldp] = acc[mi][ni][v];
      }
}

hipError_t rbrp_launch_gemm_tn(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                               int64_t K, int nsplit, int64_t kchunk, MatRef A, MatRef B,
                               double* P, int64_t ldp, int64_t pstride) {
  if (Mmax <= 0 || Nmax <= 0 || nsplit <= 0) return hipSuccess;
  dim3 grid((unsigned)((Nmax + 63) / 64), (unsigned)((Mmax + 63) / 64), (unsigned)nsplit);
  hipLaunchKernelGGL(k_gemm_tn, grid, dim3(256), 0, st, M, N, K, kchunk, A, B, P, ldp, pstride);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split-K TN for the projection W = Q_b^T A_rem and the re-orthogonalisation
// coefficients: P[z](r, c) = sum_{kk in chunk z} A(kk, r) B(kk, c) for
// column-major operands (MatRef: plain, offset or index-listed columns).
//
// Workgroup tile: 64 rows x 128 columns (an XCD-aware task); wave w owns columns
// 32w .. 32w+31 (two 16-column MFMA blocks) against the four 16-row blocks of
// A.  Each 16-row step of both operands is staged through double-buffered
// LDS ([column][row], one 128-byte run per column and step; the eight 16-byte
// units of a column are XOR-swizzled with the column index mod 8, so that the
// eight lanes of each 128-byte pass of a ds_read_b128 / ds_write_b128 touch
// eight distinct 16-byte bank groups): the next step's global
// loads are issued into staging registers before the current step's MFMAs and
// written to LDS after them (one barrier per step).  Lane (j, g) reads rows
// 4g .. 4g+3 of its column as two ds_read_b128 and MFMA q of the step consumes
// element q, so the four MFMAs of a block cover the 16 rows in the order
// 4g + q.  The MFMA computes the transposed block (B's columns as its rows):
// lane l holds P(r0 + l%16, c0 + 4v + l/16), and each partial store
// instruction writes four 128-byte runs of P.  K and kchunk are multiples of
// 16 and both leading dimensions are even; rows >= the true height are zero.
// ---------------------------------------------------------------------------
#define TN64_KS 16
// LDS offset (doubles) of 16-byte unit u (rows 2u, 2u+1) of staged column col
#define TN64_AT(col, u) ((col) * TN64_KS + 2 * ((u) ^ ((col) & 7)))

// A zero word that masked loads read instead of an address past a column's end.
__device__ double rbrp_zero_word[2];

// Address of row `row` of column `col` (true height `h`), or of the zero word
// past it: the select is on the address, so the load itself is unconditional.
__device__ __forceinline__ const double* first_src(const double* col, int64_t row, int64_t h) {
  return row < h ? col + row : rbrp_zero_word;
}

// With FIRST, B is the caller's matrix of the FP64 loop's first block (no
// working copy yet): column c of B is fs column fs.cols[fs.coff + *fs.coff_dyn
// + c] and its rows >= fs.mtrue read as zero (B itself is unused).
template <bool FIRST>
__global__ __launch_bounds__(256, 2) void k_gemm_tn64(DimSpec dM, DimSpec dN, int64_t K,
                                                      int64_t kchunk, MatRef A, MatRef B,
                                                      double* __restrict__ P, int64_t ldp,
                                                      int64_t pstride, int gx, int gy, int gz,
                                                      FirstSrc fs) {
  // task = N tile + gx * (M tile + gy * split), XCD-aware (the tiles of one
  // split share its rows of A through one L2)
  const int64_t task = rbrp_xcd_task(blockIdx.x, (int64_t)gx * gy * gz);
  if (task >= (int64_t)gx * gy * gz) return;
  const int64_t bx = task % gx, by = (task / gx) % gy, bz = task / ((int64_t)gx * gy);
  const int64_t M = dim_eval(dM), N = dim_eval(dN);
  const int64_t n0 = bx * 128, m0 = by * 64;
  if (m0 >= M || n0 >= N) return;  // uniform per workgroup
  const int64_t kb0 = bz * kchunk;
  const int64_t kb1 = min(K, kb0 + kchunk);
  if (kb0 >= kb1) return;  // empty chunk (uniform); its partials are never read

  __shared__ double As[2][64 * TN64_KS];   // [buf][col][row], swizzled
  __shared__ double Bs[2][128 * TN64_KS];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lj = lane & 15, lg = lane >> 4;
  // Stagers.  Out-of-range columns are clamped to the last valid one: they
  // only feed output rows / columns that are never stored.
  const int ac = tid >> 2, aq = (tid & 3) * 4;   // A: column ac, rows aq .. aq+3
  const int bc = tid >> 1, bq = (tid & 1) * 8;   // B: column bc, rows bq .. bq+7
  const double* ap = mat_col(A, min(m0 + ac, M - 1)) + aq;
  const double* bp;
  if (FIRST)
    bp = fs.p + (int64_t)fs.cols[fs.coff + (fs.coff_dyn ? (int64_t)*fs.coff_dyn : 0) +
                                 min(n0 + bc, N - 1)] * fs.ld;
  else
    bp = mat_col(B, min(n0 + bc, N - 1)) + bq;
  const int au = aq >> 1, bu = bq >> 1;  // first 16-byte unit of the stager
  // row blocks holding rows < M (the kept count is often below the tile height)
  const int mba = (int)min((int64_t)4, (M - m0 + 15) / 16);

  v4d acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

  // Two sets of staging registers (named scalars: small arrays here get
  // placed in LDS): the loads of step k+2 are issued before the MFMAs of step
  // k, and step k+1's registers are written to LDS after them.
#define TN64_LOAD(S, kk)                                                              \
  do {                                                                                \
    S##a0 = *(const double2*)(ap + (kk));                                             \
    S##a1 = *(const double2*)(ap + (kk) + 2);                                         \
    if (FIRST) {                                                                      \
      const int64_t tl_r = (kk) + bq;                                                 \
      S##b0 = make_double2(*first_src(bp, tl_r, fs.mtrue), *first_src(bp, tl_r + 1, fs.mtrue)); \
      S##b1 = make_double2(*first_src(bp, tl_r + 2, fs.mtrue), *first_src(bp, tl_r + 3, fs.mtrue)); \
      S##b2 = make_double2(*first_src(bp, tl_r + 4, fs.mtrue), *first_src(bp, tl_r + 5, fs.mtrue)); \
      S##b3 = make_double2(*first_src(bp, tl_r + 6, fs.mtrue), *first_src(bp, tl_r + 7, fs.mtrue)); \
    } else {                                                                          \
      S##b0 = *(const double2*)(bp + (kk));                                           \
      S##b1 = *(const double2*)(bp + (kk) + 2);                                       \
      S##b2 = *(const double2*)(bp + (kk) + 4);                                       \
      S##b3 = *(const double2*)(bp + (kk) + 6);                                       \
    }                                                                                 \
  } while (0)
#define TN64_STORE(S, bb)                                 \
  do {                                                    \
    *(double2*)&As[bb][TN64_AT(ac, au)] = S##a0;          \
    *(double2*)&As[bb][TN64_AT(ac, au + 1)] = S##a1;      \
    *(double2*)&Bs[bb][TN64_AT(bc, bu)] = S##b0;          \
    *(double2*)&Bs[bb][TN64_AT(bc, bu + 1)] = S##b1;      \
    *(double2*)&Bs[bb][TN64_AT(bc, bu + 2)] = S##b2;      \
    *(double2*)&Bs[bb][TN64_AT(bc, bu + 3)] = S##b3;      \
  } while (0)
#define TN64_STEP(bb)                                                                           \
  do {                                                                                          \
    const double* as = &As[bb][0];                                                              \
    const double* bs = &Bs[bb][0];                                                              \
    const double2 x00 = *(const double2*)&bs[TN64_AT(32 * wid + lj, 2 * lg)];                   \
    const double2 x01 = *(const double2*)&bs[TN64_AT(32 * wid + lj, 2 * lg + 1)];               \
    const double2 x10 = *(const double2*)&bs[TN64_AT(32 * wid + 16 + lj, 2 * lg)];              \
    const double2 x11 = *(const double2*)&bs[TN64_AT(32 * wid + 16 + lj, 2 * lg + 1)];          \
    _Pragma("unroll") for (int mb = 0; mb < 4; ++mb) {                                          \
      if (mb >= mba) break; /* uniform */                                                       \
      const double2 a0 = *(const double2*)&as[TN64_AT(16 * mb + lj, 2 * lg)];                   \
      const double2 a1 = *(const double2*)&as[TN64_AT(16 * mb + lj, 2 * lg + 1)];               \
      acc[mb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x00.x, a0.x, acc[mb][0], 0, 0, 0);     \
      acc[mb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x10.x, a0.x, acc[mb][1], 0, 0, 0);     \
      acc[mb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x00.y, a0.y, acc[mb][0], 0, 0, 0);     \
      acc[mb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x10.y, a0.y, acc[mb][1], 0, 0, 0);     \
      acc[mb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x01.x, a1.x, acc[mb][0], 0, 0, 0);     \
      acc[mb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x11.x, a1.x, acc[mb][1], 0, 0, 0);     \
      acc[mb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x01.y, a1.y, acc[mb][0], 0, 0, 0);     \
      acc[mb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x11.y, a1.y, acc[mb][1], 0, 0, 0);     \
    }                                                                                           \
  } while (0)

  // The loads and the LDS stores are unconditional (a step beyond the chunk
  // re-reads its last step; a store past the last step writes the buffer that
  // is not read), so that on every path the vector-memory counter waits of the
  // stores count only the loads issued before them.
  const int64_t klast = kb1 - TN64_KS;
  double2 Xa0, Xa1, Xb0, Xb1, Xb2, Xb3;
  double2 Ya0, Ya1, Yb0, Yb1, Yb2, Yb3;
  TN64_LOAD(X, kb0);
  TN64_STORE(X, 0);
  TN64_LOAD(X, min(kb0 + TN64_KS, klast));
  __syncthreads();
  // Whole pairs of steps, then an odd last step: no branch inside the loop
  // body, so no path reaches the loop head with a load still unstored.  Step k
  // reads buffer k & 1; X holds step k+1 at even k, Y at odd k.
  const int64_t nsteps = (kb1 - kb0) / TN64_KS;
  int64_t kb = kb0;
  for (int64_t s2 = 2; s2 <= nsteps; s2 += 2, kb += 2 * TN64_KS) {
    TN64_LOAD(Y, min(kb + 2 * TN64_KS, klast));
    __builtin_amdgcn_sched_barrier(0);
    TN64_STEP(0);
    __builtin_amdgcn_sched_barrier(0);
    TN64_STORE(X, 1);
    __syncthreads();
    TN64_LOAD(X, min(kb + 3 * TN64_KS, klast));
    __builtin_amdgcn_sched_barrier(0);
    TN64_STEP(1);
    __builtin_amdgcn_sched_barrier(0);
    TN64_STORE(Y, 0);
    __syncthreads();
  }
  if (nsteps & 1) TN64_STEP(0);
#undef TN64_LOAD
#undef TN64_STORE
#undef TN64_STEP
  double* Pz = P + bz * pstride;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = m0 + 16 * mb + lj;
        const int64_t c = n0 + 32 * wid + 16 * h + 4 * v + lg;
        if (r < M && c < N) Pz[r + c * ldp] = acc[mb][h][v];
      }
}

// ---------------------------------------------------------------------------
// Split-K Gram matrix of a tall block of at most 16 NBK columns (NBK = 4: 64,
// NBK = 8: 128): P[z](r, c) = sum_{kk in chunk z} A(kk, r) A(kk, c) for r, c < N.
//
// The result is cut into 16 x 16 blocks; the NBK (NBK + 1) / 2 on and above the
// diagonal are computed, consecutive runs of them (row by row) per wave
// (gram_block), blocks of columns >= N skipped, and each off-diagonal block is
// also stored transposed, so the partials hold the whole (exactly symmetric)
// matrix.  A 16-row step of A is staged once, in double-buffered LDS laid out
// as k_gemm_tn64's ([column][row], the 16-byte units of a column
// XOR-swizzled), with the loads of the step after next in flight; lane (j, g)
// reads rows 4g .. 4g+3 of a column as two ds_read_b128 and MFMA q of a block
// takes element q of both operands.  K and kchunk are multiples of 16 and the
// leading dimension is even.
// ---------------------------------------------------------------------------
template <int NB>
struct GramShape {
  static constexpr int NBLK = NB * (NB + 1) / 2;  // blocks on and above the diagonal
  static constexpr int PER = (NBLK + 3) / 4;      // blocks per wave
};
// Block k of the upper triangle in row order, (I, J) packed as 8 I + J.
__host__ __device__ constexpr int gram_upper(int k, int nb) {
  int I = 0;
  while (k >= nb - I) {
    k -= nb - I;
    ++I;
  }
  return 8 * I + I + k;
}
// Block b of wave w for an NB x NB block grid, or -1.
template <int NB>
__host__ __device__ constexpr int gram_block(int w, int b) {
  return GramShape<NB>::PER * w + b < GramShape<NB>::NBLK
             ? gram_upper(GramShape<NB>::PER * w + b, NB)
             : -1;
}

template <int NB, int W>
__device__ __forceinline__ void gram_step(const double* as, int lj, int lg,
                                          v4d (&acc)[GramShape<NB>::PER]) {
#pragma unroll
  for (int b = 0; b < GramShape<NB>::PER; ++b) {
    const int blk = gram_block<NB>(W, b);
    if (blk < 0) continue;
    const int I = blk >> 3, J = blk & 7;
    const double2 a0 = *(const double2*)&as[TN64_AT(16 * I + lj, 2 * lg)];
    const double2 a1 = *(const double2*)&as[TN64_AT(16 * I + lj, 2 * lg + 1)];
    double2 b0 = a0, b1 = a1;
    if (J != I) {
      b0 = *(const double2*)&as[TN64_AT(16 * J + lj, 2 * lg)];
      b1 = *(const double2*)&as[TN64_AT(16 * J + lj, 2 * lg + 1)];
    }
    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.x, b0.x, acc[b], 0, 0, 0);
    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, b0.y, acc[b], 0, 0, 0);
    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.x, b1.x, acc[b], 0, 0, 0);
    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, b1.y, acc[b], 0, 0, 0);
  }
}

// lane (j, g) holds G(16 I + 4 v + g, 16 J + j) of block b in acc[b][v]
template <int NB, int W>
__device__ __forceinline__ void gram_store(const v4d (&acc)[GramShape<NB>::PER], int lj, int lg,
                                           int64_t N, double* __restrict__ Pz, int64_t ldp) {
#pragma unroll
  for (int b = 0; b < GramShape<NB>::PER; ++b) {
    const int blk = gram_block<NB>(W, b);
    if (blk < 0) continue;
    const int I = blk >> 3, J = blk & 7;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = 16 * I + 4 * v + lg, c = 16 * J + lj;
      if (r < N && c < N) {
        Pz[r + c * ldp] = acc[b][v];
        if (I != J) Pz[c + r * ldp] = acc[b][v];
      }
    }
  }
}

// The K loop and the stores of one workgroup for NB column blocks (a compile-
// time count, so that each wave's blocks are straight-line code); the stage
// holds NBK blocks of columns (columns >= N re-read the last one).
template <int NBK, int NB>
__device__ __forceinline__ void gram_chunk(int64_t N, int64_t kb0, int64_t kb1, const double* ap,
                                           int ac, int au, double (*As)[16 * NBK * TN64_KS],
                                           double* __restrict__ Pz, int64_t ldp) {
  constexpr int PT = 16 * NBK * 16 / 2 / 256;  // 16-byte units staged per thread: 2 or 4
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lj = lane & 15, lg = lane >> 4;
  v4d acc[GramShape<NB>::PER];
#pragma unroll
  for (int b = 0; b < GramShape<NB>::PER; ++b) acc[b] = (v4d){0.0, 0.0, 0.0, 0.0};
  // (named staging registers: small arrays here go to scratch)
#define G_LOAD(S_, kk)                                     \
  do {                                                     \
    S_##0 = *(const double2*)(ap + (kk));                  \
    S_##1 = *(const double2*)(ap + (kk) + 2);              \
    if (PT > 2) {                                          \
      S_##2 = *(const double2*)(ap + (kk) + 4);            \
      S_##3 = *(const double2*)(ap + (kk) + 6);            \
    }                                                      \
  } while (0)
#define G_STORE(S_, bb)                                    \
  do {                                                     \
    *(double2*)&As[bb][TN64_AT(ac, au)] = S_##0;           \
    *(double2*)&As[bb][TN64_AT(ac, au + 1)] = S_##1;       \
    if (PT > 2) {                                          \
      *(double2*)&As[bb][TN64_AT(ac, au + 2)] = S_##2;     \
      *(double2*)&As[bb][TN64_AT(ac, au + 3)] = S_##3;     \
    }                                                      \
  } while (0)
#define G_STEP(bb)                                                    \
  do {                                                                \
    switch (wid) {                                                    \
      case 0: gram_step<NB, 0>(&As[bb][0], lj, lg, acc); break;       \
      case 1: gram_step<NB, 1>(&As[bb][0], lj, lg, acc); break;       \
      case 2: gram_step<NB, 2>(&As[bb][0], lj, lg, acc); break;       \
      default: gram_step<NB, 3>(&As[bb][0], lj, lg, acc); break;      \
    }                                                                 \
  } while (0)
  // the same step schedule as k_gemm_tn64: unconditional loads (a step past
  // the chunk re-reads its last step), one barrier per step
  const int64_t klast = kb1 - TN64_KS;
  double2 X0, X1, X2, X3, Y0, Y1, Y2, Y3;
  G_LOAD(X, kb0);
  G_STORE(X, 0);
  G_LOAD(X, min(kb0 + TN64_KS, klast));
  __syncthreads();
  const int64_t nsteps = (kb1 - kb0) / TN64_KS;
  int64_t kb = kb0;
  for (int64_t s2 = 2; s2 <= nsteps; s2 += 2, kb += 2 * TN64_KS) {
    G_LOAD(Y, min(kb + 2 * TN64_KS, klast));
    __builtin_amdgcn_sched_barrier(0);
    G_STEP(0);
    __builtin_amdgcn_sched_barrier(0);
    G_STORE(X, 1);
    __syncthreads();
    G_LOAD(X, min(kb + 3 * TN64_KS, klast));
    __builtin_amdgcn_sched_barrier(0);
    G_STEP(1);
    __builtin_amdgcn_sched_barrier(0);
    G_STORE(Y, 0);
    __syncthreads();
  }
  if (nsteps & 1) G_STEP(0);
#undef G_LOAD
#undef G_STORE
#undef G_STEP
  switch (wid) {
    case 0: gram_store<NB, 0>(acc, lj, lg, N, Pz, ldp); break;
    case 1: gram_store<NB, 1>(acc, lj, lg, N, Pz, ldp); break;
    case 2: gram_store<NB, 2>(acc, lj, lg, N, Pz, ldp); break;
    default: gram_store<NB, 3>(acc, lj, lg, N, Pz, ldp); break;
  }
}

template <int NBK>
__global__ __launch_bounds__(256, 2) void k_gram(DimSpec dN, int64_t K, int64_t kchunk, MatRef A,
                                                 double* __restrict__ P, int64_t ldp,
                                                 int64_t pstride) {
  const int64_t N = dim_eval(dN);
  const int64_t kb0 = (int64_t)blockIdx.x * kchunk;
  const int64_t kb1 = min(K, kb0 + kchunk);
  if (N <= 0 || kb0 >= kb1) return;  // uniform per workgroup
  __shared__ double As[2][16 * NBK * TN64_KS];
  // stager: column ac (clamped to the last one: it only feeds outputs that
  // are not stored), 2 or 4 16-byte units from row aq on
  constexpr int TPC = 256 / (16 * NBK);  // threads per column
  const int tid = threadIdx.x;
  const int ac = tid / TPC, aq = (tid % TPC) * (16 / TPC), au = aq >> 1;
  const double* ap = mat_col(A, min((int64_t)ac, N - 1)) + aq;
  double* Pz = P + (int64_t)blockIdx.x * pstride;
  const int nbk = (int)min((int64_t)NBK, (N + 15) / 16);
  if (NBK == 4) {
    switch (nbk) {
      case 1: gram_chunk<NBK, 1>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 2: gram_chunk<NBK, 2>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 3: gram_chunk<NBK, 3>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      default: gram_chunk<NBK, 4>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
    }
  } else {
    switch (nbk) {
      case 1: gram_chunk<NBK, 1>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 2: gram_chunk<NBK, 2>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 3: gram_chunk<NBK, 3>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 4: gram_chunk<NBK, 4>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 5: gram_chunk<NBK, 5>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 6: gram_chunk<NBK, 6>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      case 7: gram_chunk<NBK, 7>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
      default: gram_chunk<NBK, 8>(N, kb0, kb1, ap, ac, au, As, Pz, ldp); break;
    }
  }
}

hipError_t rbrp_launch_gram(hipStream_t st, DimSpec N, int64_t Nmax, int64_t K, int nsplit,
                            int64_t kchunk, MatRef A, double* P, int64_t ldp, int64_t pstride) {
  if (Nmax <= 0 || nsplit <= 0) return hipSuccess;
  if (Nmax > 128 || K % TN64_KS != 0 || kchunk % TN64_KS != 0 || kchunk <= 0 || A.ld % 2 != 0 ||
      ((uintptr_t)A.p & 15) || (!A.cols && A.coff_dyn) || (!A.cols && A.coff % 2 != 0) ||
      ldp < Nmax)
    return hipErrorInvalidValue;
  if (Nmax <= 64)
    hipLaunchKernelGGL(k_gram<4>, dim3((unsigned)nsplit), dim3(256), 0, st, N, K, kchunk, A, P, ldp,
                       pstride);
  else
    hipLaunchKernelGGL(k_gram<8>, dim3((unsigned)nsplit), dim3(256), 0, st, N, K, kchunk, A, P, ldp,
                       pstride);
  return hipGetLastError();
}

int rbrp_gemm_tn64_resident() {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_gemm_tn64<false>, 256, 0) != hipSuccess)
    return 0;
  return cus * per;
}

hipError_t rbrp_launch_gemm_tn64(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                                 int64_t K, int nsplit, int64_t kchunk, MatRef A, MatRef B, double* P,
                                 int64_t ldp, int64_t pstride, hipEvent_t e0, hipEvent_t e1,
                                 const FirstSrc* first) {
  if (Mmax <= 0 || Nmax <= 0 || nsplit <= 0) return hipSuccess;
  if (kchunk % TN64_KS != 0 || K % TN64_KS != 0 || A.ld % 2 != 0 || (!first && B.ld % 2 != 0))
    return hipErrorInvalidValue;
  if (first && (!first->p || !first->cols)) return hipErrorInvalidValue;
  const int64_t gx = (Nmax + 127) / 128, gy = (Mmax + 63) / 64;
  dim3 grid((unsigned)rbrp_xcd_grid(gx * gy * nsplit));
  const FirstSrc fs = first ? *first : FirstSrc{nullptr, 0, nullptr, 0, nullptr, 0};
  if (first)
    RBRP_LAUNCH_EV(e0, e1, k_gemm_tn64<true>, grid, dim3(256), 0, st, M, N, K, kchunk, A, B, P, ldp,
                   pstride, (int)gx, (int)gy, nsplit, fs);
  else
    RBRP_LAUNCH_EV(e0, e1, k_gemm_tn64<false>, grid, dim3(256), 0, st, M, N, K, kchunk, A, B, P, ldp,
                   pstride, (int)gx, (int)gy, nsplit, fs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tall NN with a long reduction (the re-orthogonalisation V -= Q C):
// C(r, c) (op)= sum_{k < K} A(r, k) B(k, c) for r < M, c < N, any K.
//
// The rows are cut into 16-row blocks and the blocks into gr contiguous
// ranges of R <= 4 blocks (R differs by at most one between ranges); a
// workgroup takes one range against one 64-column tile (an XCD-aware task, so
// the column tiles of a range share its rows of A through one L2).  Wave w
// owns the 16-column block w of the tile against the R row blocks, so a
// range of R blocks puts R units of MFMA work on each SIMD: with gr sized to
// the resident workgroups, every SIMD carries about the same number of row
// blocks whatever M is (64-row tiles would leave some compute units one tile
// more than the rest).  Each 16-deep step of both operands is staged through
// double-buffered LDS with two steps of global loads in flight: A as
// [row][k] (16 k per row), B as [column][k], the eight 16-byte units of a row
// / column XOR-swizzled with its index mod 8 (the eight lanes of each 128-byte
// LDS pass touch distinct bank groups).  Lane (j, g) reads k = 4g .. 4g+3 of
// its row / column as two ds_read_b128 and MFMA q of the step consumes element
// q.  The MFMA computes the transposed block (B's columns as its rows): lane l
// holds C(r0 + 16 rb + l%16, c0 + 16w + 4v + l/16), so every C load / store
// instruction moves four 128-byte column runs.  Reduction indices >= K read as
// zero (clamped loads, then a bit mask on B).  A may be a plain, offset or
// index-listed MatRef (the column address of an A load is uniform per wave).
// ---------------------------------------------------------------------------
#define NN64_KS 16
#define NN64_RB 4  // row blocks per workgroup at most
#define NN64_AT(idx, u) ((idx) * NN64_KS + 2 * ((u) ^ ((idx) & 7)))
// x with its bits and-ed with the mask m (0 or all ones): a loaded value kept
// or replaced by +0 without a branch, so that every load stays unconditional
__device__ __forceinline__ double nn64_mask(double x, long long m) {
  return __longlong_as_double(__double_as_longlong(x) & m);
}

// The K loop of one workgroup for R row blocks (R a template argument so that
// a range of R blocks issues exactly R blocks of MFMAs).
// Element `aoff` of column c of A: FP64, or (A32) an FP32 matrix held in
// A.p as a float pointer, widened.
template <bool A32>
__device__ __forceinline__ double nn64_a(const MatRef& A, int64_t c, uint32_t aoff) {
  if (A32) {
    const int64_t cc =
        A.cols ? (int64_t)A.cols[c] : (A.coff + (A.coff_dyn ? (int64_t)*A.coff_dyn : 0) + c);
    return (double)((const float*)A.p)[cc * A.ld + aoff];
  }
  return mat_col(A, c)[aoff];
}

template <int R, bool A32 = false>
__device__ __forceinline__ void nn64_loop(int64_t K, const MatRef& A, uint32_t aoff,
                                          const double* __restrict__ bcol, double* As0,
                                          double* Bs0, v4d* acc) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform per wave
  const int lj = lane & 15, lg = lane >> 4;
  // A stager: row ar, k = ak .. ak+3 (ak is uniform per wave, so each A load is
  // a scalar column address plus the lane's 32-bit row offset aoff from the
  // workgroup's first row)
  const int ar = tid & 63, ak = wid * 4;
  const int bc = tid >> 2, bk = (tid & 3) * 4;   // B stager: column bc, k = bk .. bk+3
  double* As[2] = {As0, As0 + 64 * NN64_KS};
  double* Bs[2] = {Bs0, Bs0 + 64 * NN64_KS};

// (the macro's locals carry an nl_ prefix: a local named like a variable of
// the argument expression would be initialised from itself)
#define NN64_LOAD(S, kk)                                                                 \
  do {                                                                                   \
    const int64_t nl_ka = (kk) + ak;                                                     \
    S##a0 = make_double2(nn64_a<A32>(A, min(nl_ka, K - 1), aoff),                        \
                         nn64_a<A32>(A, min(nl_ka + 1, K - 1), aoff));                   \
    S##a1 = make_double2(nn64_a<A32>(A, min(nl_ka + 2, K - 1), aoff),                    \
                         nn64_a<A32>(A, min(nl_ka + 3, K - 1), aoff));                   \
    const int64_t nl_kb = (kk) + bk;                                                     \
    S##b0 = make_double2(bcol[min(nl_kb, K - 1)], bcol[min(nl_kb + 1, K - 1)]);          \
    S##b1 = make_double2(bcol[min(nl_kb + 2, K - 1)], bcol[min(nl_kb + 3, K - 1)]);      \
  } while (0)
// B's rows >= K are zeroed here, at the store (a mask applied where the loads
// are issued would make the compiler wait for them there).  A's clamped
// columns are not masked: they only meet those zero rows of B.
#define NN64_STORE(S, bb, kk)                                                              \
  do {                                                                                     \
    const int64_t nl_kb = (kk) + bk;                                                       \
    *(double2*)&As[bb][NN64_AT(ar, ak >> 1)] = S##a0;                                      \
    *(double2*)&As[bb][NN64_AT(ar, (ak >> 1) + 1)] = S##a1;                                \
    *(double2*)&Bs[bb][NN64_AT(bc, bk >> 1)] =                                             \
        make_double2(nn64_mask(S##b0.x, nl_kb < K ? -1ll : 0ll),                           \
                     nn64_mask(S##b0.y, nl_kb + 1 < K ? -1ll : 0ll));                      \
    *(double2*)&Bs[bb][NN64_AT(bc, (bk >> 1) + 1)] =                                       \
        make_double2(nn64_mask(S##b1.x, nl_kb + 2 < K ? -1ll : 0ll),                       \
                     nn64_mask(S##b1.y, nl_kb + 3 < K ? -1ll : 0ll));                      \
  } while (0)
#define NN64_STEP(bb)                                                                      \
  do {                                                                                     \
    const double* as = As[bb];                                                             \
    const double* bs = Bs[bb];                                                             \
    const double2 x0 = *(const double2*)&bs[NN64_AT(16 * wid + lj, 2 * lg)];               \
    const double2 x1 = *(const double2*)&bs[NN64_AT(16 * wid + lj, 2 * lg + 1)];           \
    _Pragma("unroll") for (int rb = 0; rb < R; ++rb) {                                     \
      const double2 y0 = *(const double2*)&as[NN64_AT(16 * rb + lj, 2 * lg)];             \
      const double2 y1 = *(const double2*)&as[NN64_AT(16 * rb + lj, 2 * lg + 1)];         \
      acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0.x, y0.x, acc[rb], 0, 0, 0);        \
      acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0.y, y0.y, acc[rb], 0, 0, 0);        \
      acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.x, y1.x, acc[rb], 0, 0, 0);        \
      acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.y, y1.y, acc[rb], 0, 0, 0);        \
    }                                                                                      \
  } while (0)

  // two steps of loads in flight; the loads are unconditional (beyond K they
  // read clamped addresses), so the vector-memory waits of the stores count
  // only the loads issued before them
  double2 Xa0, Xa1, Xb0, Xb1;
  double2 Ya0, Ya1, Yb0, Yb1;
  NN64_LOAD(X, 0);
  NN64_STORE(X, 0, 0);
  NN64_LOAD(X, NN64_KS);
  __syncthreads();
  // Whole pairs of steps, then an odd last step: no branch inside the loop
  // body, and the stores are unconditional (one past the last step writes the
  // buffer that is not read), so every path waits for its loads at the store
  // and nowhere else.
  const int64_t nsteps = (K + NN64_KS - 1) / NN64_KS;
  int64_t kb = 0;
  for (int64_t s2 = 2; s2 <= nsteps; s2 += 2, kb += 2 * NN64_KS) {
    NN64_LOAD(Y, kb + 2 * NN64_KS);
    __builtin_amdgcn_sched_barrier(0);
    NN64_STEP(0);
    __builtin_amdgcn_sched_barrier(0);
    NN64_STORE(X, 1, kb + NN64_KS);
    __syncthreads();
    NN64_LOAD(X, kb + 3 * NN64_KS);
    __builtin_amdgcn_sched_barrier(0);
    NN64_STEP(1);
    __builtin_amdgcn_sched_barrier(0);
    NN64_STORE(Y, 0, kb + 2 * NN64_KS);
    __syncthreads();
  }
  if (nsteps & 1) NN64_STEP(0);
#undef NN64_LOAD
#undef NN64_STORE
#undef NN64_STEP
}

// With C32, the new C is also stored in FP32 at C32[r + c * ld32] (the mixed
// loop's FP32 copy of the block's basis).
__global__ __launch_bounds__(256, 4) void k_gemm_nn64(DimSpec dM, DimSpec dN, DimSpec dK, MatRef A,
                                                      MatRef B, MatRef C, int mode, int gx,
                                                      int gr, int64_t rbt, float* __restrict__ C32,
                                                      int64_t ld32) {
  // task = column tile + gx * row range, XCD-aware
  const int64_t task = rbrp_xcd_task(blockIdx.x, (int64_t)gx * gr);
  if (task >= (int64_t)gx * gr) return;
  const int64_t g = task / gx;
  const int64_t rb0 = g * rbt / gr, rb1 = (g + 1) * rbt / gr;
  const int R = (int)(rb1 - rb0);
  const int64_t M = dim_eval(dM), N = dim_eval(dN), K = dim_eval(dK);
  const int64_t r0 = rb0 * 16, c0 = (task % gx) * 64;
  if (R <= 0 || r0 >= M || c0 >= N || K <= 0) return;  // uniform per workgroup

  __shared__ double As[2 * 64 * NN64_KS];  // [buf][row][k], swizzled
  __shared__ double Bs[2 * 64 * NN64_KS];  // [buf][col][k], swizzled

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lj = lane & 15, lg = lane >> 4;
  // A stager row: rows past the range re-read its last row (the loads stay
  // unconditional and hit the cache), rows >= M re-read row M-1; neither feeds
  // a stored output
  const int64_t arow = min(r0 + min(tid & 63, 16 * R - 1), M - 1);
  MatRef Ar = A;  // A's rows from the workgroup's first row on
  Ar.p += r0;
  const uint32_t aoff = (uint32_t)(arow - r0);
  const double* bcol = mat_col(B, min(c0 + (tid >> 2), N - 1));

  v4d acc[NN64_RB];
#pragma unroll
  for (int rb = 0; rb < NN64_RB; ++rb) acc[rb] = (v4d){0.0, 0.0, 0.0, 0.0};
  switch (R) {
    case 1: nn64_loop<1>(K, Ar, aoff, bcol, As, Bs, acc); break;
    case 2: nn64_loop<2>(K, Ar, aoff, bcol, As, Bs, acc); break;
    case 3: nn64_loop<3>(K, Ar, aoff, bcol, As, Bs, acc); break;
    default: nn64_loop<4>(K, Ar, aoff, bcol, As, Bs, acc); break;
  }
  const int64_t c = c0 + 16 * wid + lg;
#pragma unroll
  for (int rb = 0; rb < NN64_RB; ++rb) {
    const int64_t r = r0 + 16 * rb + lj;
    if (rb < R && r < M) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (c + 4 * v < N) {
          double* d = mat_col(C, c + 4 * v) + r;
          const double x = acc[rb][v];
          const double y = (mode == GEMM_SUB) ? *d - x : ((mode == GEMM_ADD) ? *d + x : x);
          *d = y;
          if (C32) C32[r + (c + 4 * v) * ld32] = (float)y;
        }
    }
  }
}

// Norms form of k_gemm_nn64 (the FP64 residual recompute of the mixed loop's
// stopping test): E = C - A B is not stored; normpart[g * ldnp + c] receives
// sum over the rows of row range g of E(r, c)^2, for r < M, c < N.  A and C
// may be the caller's matrix (FP64, or FP32 with A32: A.p / C.p hold float
// pointers) read through index-listed MatRefs; M is the true height (A's rows
// are clamped to M - 1 by the stager, C's rows >= M are never read).
template <bool A32>
__global__ __launch_bounds__(256, 4) void k_gemm_nn64_resnorm(int64_t M, int64_t N, int64_t K,
                                                              MatRef A, MatRef B, MatRef C, int gx,
                                                              int gr, int64_t rbt,
                                                              double* __restrict__ normpart,
                                                              int64_t ldnp) {
  const int64_t task = rbrp_xcd_task(blockIdx.x, (int64_t)gx * gr);
  if (task >= (int64_t)gx * gr) return;
  const int64_t g = task / gx;
  const int64_t rb0 = g * rbt / gr, rb1 = (g + 1) * rbt / gr;
  const int R = (int)(rb1 - rb0);
  const int64_t r0 = rb0 * 16, c0 = (task % gx) * 64;
  if (R <= 0 || r0 >= M || c0 >= N || K <= 0) return;  // uniform per workgroup

  __shared__ double As[2 * 64 * NN64_KS];
  __shared__ double Bs[2 * 64 * NN64_KS];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int lj = lane & 15, lg = lane >> 4;
  const int64_t arow = min(r0 + min(tid & 63, 16 * R - 1), M - 1);
  MatRef Ar = A;
  if (A32)
    Ar.p = (double*)((const float*)A.p + r0);
  else
    Ar.p += r0;
  const uint32_t aoff = (uint32_t)(arow - r0);
  const double* bcol = mat_col(B, min(c0 + (tid >> 2), N - 1));

  v4d acc[NN64_RB];
#pragma unroll
  for (int rb = 0; rb < NN64_RB; ++rb) acc[rb] = (v4d){0.0, 0.0, 0.0, 0.0};
  switch (R) {
    case 1: nn64_loop<1, A32>(K, Ar, aoff, bcol, As, Bs, acc); break;
    case 2: nn64_loop<2, A32>(K, Ar, aoff, bcol, As, Bs, acc); break;
    case 3: nn64_loop<3, A32>(K, Ar, aoff, bcol, As, Bs, acc); break;
    default: nn64_loop<4, A32>(K, Ar, aoff, bcol, As, Bs, acc); break;
  }
  const int64_t c = c0 + 16 * wid + lg;
  double sq[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int rb = 0; rb < NN64_RB; ++rb) {
    const int64_t r = r0 + 16 * rb + lj;
    if (rb < R && r < M) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (c + 4 * v < N) {
          const double cin = A32 ? (double)nn64_a<true>(C, c + 4 * v, (uint32_t)r)
                                 : mat_col(C, c + 4 * v)[r];
          const double e = cin - acc[rb][v];
          sq[v] = fma(e, e, sq[v]);
        }
    }
  }
  // the sixteen lanes of one column (lj) added in a fixed order
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double x = sq[v];
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 8, 64);
    if (lj == 0 && c + 4 * v < N) normpart[g * ldnp + c + 4 * v] = x;
  }
}

// Row ranges of k_gemm_nn64 (and its norms form) for an Mmax x Nmax output.
int64_t rbrp_gemm_nn64_ranges(int64_t Mmax, int64_t Nmax, int resident) {
  const int64_t gx = (Nmax + 63) / 64, rbt = (Mmax + 15) / 16;
  const int64_t per = std::max<int64_t>(1, (resident > 0 ? resident : 1024) / gx);
  return std::max<int64_t>((rbt + NN64_RB - 1) / NN64_RB, std::min<int64_t>(rbt, per));
}

hipError_t rbrp_launch_gemm_nn64_resnorm(hipStream_t st, int64_t M, int64_t N, int64_t K, MatRef A,
                                         MatRef B, MatRef C, int a32, double* normpart, int64_t ldnp,
                                         int resident) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int64_t gx = (N + 63) / 64, rbt = (M + 15) / 16;
  const int64_t gr = rbrp_gemm_nn64_ranges(M, N, resident);
  if (ldnp < N) return hipErrorInvalidValue;
  dim3 grid((unsigned)rbrp_xcd_grid(gx * gr));
  if (a32)
    hipLaunchKernelGGL(k_gemm_nn64_resnorm<true>, grid, dim3(256), 0, st, M, N, K, A, B, C, (int)gx,
                       (int)gr, rbt, normpart, ldnp);
  else
    hipLaunchKernelGGL(k_gemm_nn64_resnorm<false>, grid, dim3(256), 0, st, M, N, K, A, B, C, (int)gx,
                       (int)gr, rbt, normpart, ldnp);
  return hipGetLastError();
}

int rbrp_gemm_nn64_resident() {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_gemm_nn64, 256, 0) != hipSuccess)
    return 0;
  return cus * per;
}

hipError_t rbrp_launch_gemm_nn64(hipStream_t st, DimSpec M, int64_t Mmax, DimSpec N, int64_t Nmax,
                                 DimSpec K, MatRef A, MatRef B, MatRef C, int mode, int resident,
                                 float* C32, int64_t ld32) {
  if (Mmax <= 0 || Nmax <= 0) return hipSuccess;
  const int64_t gx = (Nmax + 63) / 64, rbt = (Mmax + 15) / 16;
  // as many row ranges as resident workgroups per column tile, each of at
  // most NN64_RB row blocks
  const int64_t per = std::max<int64_t>(1, (resident > 0 ? resident : 1024) / gx);
  const int64_t gr = std::max<int64_t>((rbt + NN64_RB - 1) / NN64_RB, std::min<int64_t>(rbt, per));
  dim3 grid((unsigned)rbrp_xcd_grid(gx * gr));
  hipLaunchKernelGGL(k_gemm_nn64, grid, dim3(256), 0, st, M, N, K, A, B, C, mode, (int)gx, (int)gr,
                     rbt, C32, ld32);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split reduction: dst(r, c) (=|+=|-=) sum_{z=0}^{nsplit-1} P[z](r, c); with
// dst2 also stored row-major (dst2[r * ld2 + c]), with radd also added to
// radd[c + racols[r] * ldra] (the re-orthogonalisation coefficients into the
// sampled columns of R).
// One workgroup per (64-row slab, column): sixteen waves each sum every
// sixteenth split with eight loads in flight per lane (eight accumulators),
// then the sixteen partials are added in a fixed tree order, so the result
// does not depend on scheduling.
// ---------------------------------------------------------------------------
#define RED_GROUPS 16
// The stores of one reduced sum s of (r, c).
template <typename T2>
__device__ __forceinline__ void reduce_store(double s, int64_t r, int64_t c, const MatRef& dst,
                                             int mode, T2* __restrict__ dst2, int64_t ld2,
                                             double* __restrict__ radd, int64_t ldra,
                                             const int* __restrict__ racols) {
  if (dst.p) {
    double* d = mat_col(dst, c) + r;
    if (mode == GEMM_ADD)
      *d += s;
    else if (mode == GEMM_SUB)
      *d -= s;
    else
      *d = s;
  }
  if (dst2) dst2[r * ld2 + c] = (T2)s;  // row-major copy
  if (radd) radd[c + (int64_t)racols[r] * ldra] += s;
}

// Few splits (at most RED_FEW): one thread per (r, c), four columns of 64 rows
// per workgroup; split z goes to accumulator z % 8, the eight added in a fixed
// tree, all loads of a group of eight splits in flight.
#define RED_FEW 32
template <typename T2>
__global__ __launch_bounds__(256) void k_reduce_few(
    DimSpec dM, DimSpec dN, const double* __restrict__ P, int64_t ldp, int64_t pstride, int nsplit,
    MatRef dst, int mode, T2* __restrict__ dst2, int64_t ld2, double* __restrict__ radd,
    int64_t ldra, const int* __restrict__ racols) {
  const int64_t M = dim_eval(dM), N = dim_eval(dN);
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t r = (int64_t)blockIdx.y * 64 + (threadIdx.x & 63);
  if (c >= N || r >= M) return;
  const double* p = P + r + c * ldp;
  double acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = 0.0;
  for (int z0 = 0; z0 < nsplit; z0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = (z0 + u < nsplit) ? p[(int64_t)(z0 + u) * pstride] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += v[u];
  }
  const double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  reduce_store<T2>(s, r, c, dst, mode, dst2, ld2, radd, ldra, racols);
}

template <typename T2>
__global__ __launch_bounds__(64 * RED_GROUPS) void k_reduce_splits(
    DimSpec dM, DimSpec dN, const double* __restrict__ P, int64_t ldp, int64_t pstride, int nsplit,
    MatRef dst, int mode, T2* __restrict__ dst2, int64_t ld2, double* __restrict__ radd,
    int64_t ldra, const int* __restrict__ racols) {
  __shared__ double part[RED_GROUPS][64];
  const int64_t M = dim_eval(dM), N = dim_eval(dN);
  const int64_t c = blockIdx.x;
  const int lr = threadIdx.x & 63, zg = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.y * 64 + lr;
  if (c >= N || (int64_t)blockIdx.y * 64 >= M) return;  // uniform per workgroup
  double acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = 0.0;
  if (r < M) {
    const double* p = P + r + c * ldp;
    int z = zg;
    for (; z + 7 * RED_GROUPS < nsplit; z += 8 * RED_GROUPS) {
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += p[(int64_t)(z + u * RED_GROUPS) * pstride];
    }
    for (; z < nsplit; z += RED_GROUPS) acc[0] += p[(int64_t)z * pstride];
  }
  part[zg][lr] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  __syncthreads();
  if (zg == 0 && r < M) {
    double t[RED_GROUPS / 2];
#pragma unroll
    for (int g = 0; g < RED_GROUPS / 2; ++g) t[g] = part[2 * g][lr] + part[2 * g + 1][lr];
#pragma unroll
    for (int w = RED_GROUPS / 4; w >= 1; w >>= 1)
#pragma unroll
      for (int g = 0; g < w; ++g) t[g] = t[2 * g] + t[2 * g + 1];
    reduce_store<T2>(t[0], r, c, dst, mode, dst2, ld2, radd, ldra, racols);
  }
}

hipError_t rbrp_launch_reduce_splits(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                                     const double* P, int64_t ldp, int64_t pstride, int nsplit,
                                     MatRef dst, int mode, double* dst2, int64_t ld2, double* radd,
                                     int64_t ldra, const int* racols) {
  if (Mmax <= 0 || Nmax <= 0) return hipSuccess;
  if (radd && !racols) return hipErrorInvalidValue;
  if (nsplit <= RED_FEW)
    hipLaunchKernelGGL(k_reduce_few<double>, dim3((unsigned)((Nmax + 3) / 4), (unsigned)((Mmax + 63) / 64)),
                       dim3(256), 0, st, M, N, P, ldp, pstride, nsplit, dst, mode, dst2, ld2, radd,
                       ldra, racols);
  else
    hipLaunchKernelGGL(k_reduce_splits<double>, dim3((unsigned)Nmax, (unsigned)((Mmax + 63) / 64)),
                       dim3(64 * RED_GROUPS), 0, st, M, N, P, ldp, pstride, nsplit, dst, mode, dst2,
                       ld2, radd, ldra, racols);
  return hipGetLastError();
}

hipError_t rbrp_launch_reduce_splits_f32copy(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax,
                                             int64_t Nmax, const double* P, int64_t ldp,
                                             int64_t pstride, int nsplit, MatRef dst, int mode,
                                             float* dst2, int64_t ld2) {
  if (Mmax <= 0 || Nmax <= 0) return hipSuccess;
  if (nsplit <= RED_FEW)
    hipLaunchKernelGGL(k_reduce_few<float>, dim3((unsigned)((Nmax + 3) / 4), (unsigned)((Mmax + 63) / 64)),
                       dim3(256), 0, st, M, N, P, ldp, pstride, nsplit, dst, mode, dst2, ld2,
                       (double*)nullptr, (int64_t)0, (const int*)nullptr);
  else
    hipLaunchKernelGGL(k_reduce_splits<float>, dim3((unsigned)Nmax, (unsigned)((Mmax + 63) / 64)),
                       dim3(64 * RED_GROUPS), 0, st, M, N, P, ldp, pstride, nsplit, dst, mode, dst2,
                       ld2, (double*)nullptr, (int64_t)0, (const int*)nullptr);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// NN with update: C(r, c) = (mode) A(r, 0:K) B(0:K, c); optional per-tile
// squared column norms of the result.  The accumulator tile is staged through
// LDS so that C is read and written in 512-byte column runs; each wave issues
// the loads of all 16 of its C columns before using any of them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm_nn(DimSpec dM, DimSpec dN, DimSpec dK, MatRef A,
                                                 MatRef B, MatRef C, int mode,
                                                 double* __restrict__ normpart, int64_t ldnp) {
  const int64_t Mr = dim_eval(dM), N = dim_eval(dN), K = dim_eval(dK);
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  if (r0 >= Mr || c0 >= N) return;

  __shared__ double As[KT * 65];         // [k][r]
  __shared__ double Bs[64 * (KT + 1)];   // [c][k]
  __shared__ double Cs[64 * 65];         // [c][r]

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  // A loader: k row ak, rows r0 + arq .. +3
  const int ak = tid >> 4, arq = (tid & 15) * 4;
  // B loader: column bc, k rows bkq .. +3
  const int bc = tid >> 2, bkq = (tid & 3) * 4;
  const double* bp = (c0 + bc < N) ? mat_col(B, c0 + bc) : nullptr;

  double ra[4], rb[4];
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

  auto load = [&](int64_t kb) {
    const double* ap = (kb + ak < K) ? mat_col(A, kb + ak) : nullptr;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r = r0 + arq + q;
      ra[q] = (ap && r < Mr) ? ap[r] : 0.0;
      const int64_t kr = kb + bkq + q;
      rb[q] = (bp && kr < K) ? bp[kr] : 0.0;
    }
  };

  if (K > 0) load(0);
  for (int64_t kb = 0; kb < K; kb += KT) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[ak * 65 + arq + q] = ra[q];
      Bs[bc * (KT + 1) + bkq + q] = rb[q];
    }
    __syncthreads();
    if (kb + KT < K) load(kb + KT);
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4) {
      const int ko = kk + (lane >> 4);
      const double a0 = As[ko * 65 + wm * 32 + (lane & 15)];
      const double a1 = As[ko * 65 + wm * 32 + 16 + (lane & 15)];
      const double b0 = Bs[(wn * 32 + (lane & 15)) * (KT + 1) + ko];
      const double b1 = Bs[(wn * 32 + 16 + (lane & 15)) * (KT + 1) + ko];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  // wave `wid` updates columns wid, wid+4, ..., wid+60; lane = row within the tile.
  // Issue the C loads first (independent of the accumulators) ...
  const int64_t r = r0 + lane;
  const bool rok = r < Mr;
  const int ncols = (int)min((int64_t)64, N - c0);
  double cv[16];
  if (mode != GEMM_SET) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = wid + 4 * q;
      cv[q] = (c < ncols && rok) ? mat_col(C, c0 + c)[r] : 0.0;
    }
  }
  // ... then stage the accumulators as Cs[c][r]
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rr = wm * 32 + mi * 16 + mf_row(lane, v);
        const int c = wn * 32 + ni * 16 + mf_col(lane);
        Cs[c * 65 + rr] = acc[mi][ni][v];
      }
  __syncthreads();
  double nsq[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = wid + 4 * q;
    const double v = Cs[c * 65 + lane];
    double nv;
    if (mode == GEMM_SUB)
      nv = cv[q] - v;
    else if (mode == GEMM_ADD)
      nv = cv[q] + v;
    else
      nv = v;
    if (c < ncols && rok) mat_col(C, c0 + c)[r] = nv;
    nsq[q] = (c < ncols && rok) ? nv * nv : 0.0;
  }
  if (normpart) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = wid + 4 * q;
      const double t = rbrp_wave_sum(nsq[q]);
      if (lane == 0 && c < ncols) normpart[(int64_t)blockIdx.y * ldnp + c0 + c] = t;
    }
  }
}

hipError_t rbrp_launch_gemm_nn(hipStream_t st, DimSpec M, int64_t Mmax, DimSpec N, int64_t Nmax,
                               DimSpec K, MatRef A, MatRef B, MatRef C, int mode, double* normpart,
                               int64_t ldnp) {
  if (Mmax <= 0 || Nmax <= 0) return hipSuccess;
  dim3 grid((unsigned)((Nmax + 63) / 64), (unsigned)((Mmax + 63) / 64));
  hipLaunchKernelGGL(k_gemm_nn, grid, dim3(256), 0, st, M, N, K, A, B, C, mode, normpart, ldnp);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tall rank-K update (the trailing-matrix update and the CholQR2 products):
// C(r, c) (op)= sum_k A(r, k) B(k, c) for r < M, c < N, K <= 128, optional
// squared column norms of the new C, one partial per workgroup row group
// (accumulated per lane over the workgroup's sweep, reduced once at its end).
//
// The MFMA computes the transposed product D = B^T A^T so that the result
// layout puts consecutive rows of C in consecutive lanes: lane l holds
// C(row0 + l%16, col0 + 4v + l/16), i.e. every C load/store instruction moves
// four 128-byte column runs, with no LDS staging of C.
//
// Workgroup = 4 waves stacked along rows, 64 JB rows x 32 columns per tile
// (JB = 1 in the loop: 16 rows per wave keeps the register footprint low
// enough for four waves per SIMD), sweeping `T` consecutive row tiles: B (K x 32) sits in LDS for the whole
// sweep; for each tile the A fragments (16 consecutive rows of one A column
// per 16 lanes) are loaded first and the next tile's C is prefetched right
// after them, so the A loads can be waited on while the C prefetch is still
// in flight (the vector-memory counter retires loads in issue order).
// ---------------------------------------------------------------------------
// A and C are plain column-major matrices whose first column is
// coff + (*coff_dyn) (the trailing block of the working matrix); Bt is the
// row-major copy of B written by the projection's split reduction.
template <int JB, int KSTEPS, bool NORMS>
__global__ __launch_bounds__(256) void k_gemm_trail(int64_t M, DimSpec dN, DimSpec dK, int T,
                                                    const double* __restrict__ A0, int64_t lda,
                                                    int64_t acoff, const int* acoff_dyn,
                                                    const double* __restrict__ Bt, int64_t ldbt,
                                                    double* __restrict__ C0, int64_t ldc,
                                                    int64_t ccoff, const int* ccoff_dyn,
                                                    double* __restrict__ normpart, int64_t ldnp) {
  __shared__ double Ws[4 * KSTEPS * 32];  // [k][32]
  __shared__ double red[4][32];
  const int64_t N = dim_eval(dN), K = dim_eval(dK);
  const int64_t c0 = (int64_t)blockIdx.x * 32;
  const int64_t rows_per_tile = 64 * JB;
  const int64_t rbase = (int64_t)blockIdx.y * T * rows_per_tile;
  if (c0 >= N || rbase >= M) return;  // uniform per workgroup
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const double* A = A0 + (acoff + (acoff_dyn ? *acoff_dyn : 0)) * lda;
  double* C = C0 + (ccoff + (ccoff_dyn ? *ccoff_dyn : 0) + c0) * ldc;
  for (int idx = threadIdx.x; idx < 4 * KSTEPS * 32; idx += 256) {
    const int k = idx >> 5, c = idx & 31;
    Ws[idx] = (k < K && c0 + c < N) ? Bt[(int64_t)k * ldbt + c0 + c] : 0.0;
  }
  __syncthreads();
  // column offsets (within the 32-column tile) held by this lane: 16*ib + 4v + lk
  const int ncol = (int)min((int64_t)32, N - c0);

  double cvA[JB][2][4], cvB[JB][2][4];
  auto loadC = [&](int64_t r0, double (&cv)[JB][2][4]) {
    const int64_t rw = r0 + 16 * JB * wid;
#pragma unroll
    for (int jb = 0; jb < JB; ++jb)
#pragma unroll
      for (int ib = 0; ib < 2; ++ib)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int cc = 16 * ib + 4 * v + lk;
          const int64_t r = rw + 16 * jb + li;
          cv[jb][ib][v] = (cc < ncol && r < M) ? C[r + cc * ldc] : 0.0;
        }
  };
  // squared norms of the new C, accumulated per lane over the whole sweep and
  // reduced once at its end (one partial per workgroup row group)
  double nsq[2][4];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int v = 0; v < 4; ++v) nsq[ib][v] = 0.0;
  loadC(rbase, cvA);
  for (int t = 0; t < T; ++t) {
    const int64_t r0 = rbase + (int64_t)t * rows_per_tile;
    if (r0 >= M) break;  // uniform per workgroup
    const int64_t rw = r0 + 16 * JB * wid;
    double fa[KSTEPS][JB];
#pragma unroll
    for (int q = 0; q < KSTEPS; ++q) {
      const int k = 4 * q + lk;
#pragma unroll
      for (int jb = 0; jb < JB; ++jb) {
        const int64_t r = rw + 16 * jb + li;
        fa[q][jb] = (k < K && r < M) ? A[r + (int64_t)k * lda] : 0.0;
      }
    }
    if (t + 1 < T) loadC(r0 + rows_per_tile, cvB);
    v4d acc[JB][2];
#pragma unroll
    for (int a = 0; a < JB; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < KSTEPS; ++q) {
      const int k = 4 * q + lk;
      const double b0 = Ws[k * 32 + li], b1 = Ws[k * 32 + 16 + li];
#pragma unroll
      for (int jb = 0; jb < JB; ++jb) {
        acc[jb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, fa[q][jb], acc[jb][0], 0, 0, 0);
        acc[jb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, fa[q][jb], acc[jb][1], 0, 0, 0);
      }
    }
    // D[i][j] (i = column within the 16-column block, j = row) sits in lane
    // j + 16*(i%4), element i/4: with i = 4v + lk that is lane li + 16*lk.
#pragma unroll
    for (int jb = 0; jb < JB; ++jb)
#pragma unroll
      for (int ib = 0; ib < 2; ++ib)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int cc = 16 * ib + 4 * v + lk;
          const int64_t r = rw + 16 * jb + li;
          const double nv = cvA[jb][ib][v] - acc[jb][ib][v];
          if (cc < ncol && r < M) {
            C[r + cc * ldc] = nv;
            nsq[ib][v] = fma(nv, nv, nsq[ib][v]);
          }
        }
#pragma unroll
    for (int jb = 0; jb < JB; ++jb)
#pragma unroll
      for (int ib = 0; ib < 2; ++ib)
#pragma unroll
        for (int v = 0; v < 4; ++v) cvA[jb][ib][v] = cvB[jb][ib][v];
  }
  if (NORMS) {
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double x = nsq[ib][v];
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        x += __shfl_xor(x, 8, 64);
        if (li == 0) red[wid][16 * ib + 4 * v + lk] = x;
      }
    __syncthreads();
    if (wid == 0 && lane < ncol)
      normpart[(int64_t)blockIdx.y * ldnp + c0 + lane] =
          (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
  }
}

// ---------------------------------------------------------------------------
// FP64 trailing update for K <= 64 and M a multiple of 64 (the RBRP loop's
// shape): C(r, c) -= sum_{k<K} A(r, k) Bt[k * ldbt + c], optional squared column
// norms of the new C (one partial per workgroup row group).
//
// Workgroup = 4 waves side by side along columns (32 each, 128 per workgroup),
// sweeping T row tiles of 64 rows.  The 64 x 64 block of A of a row tile is
// staged once in LDS ([row][k], pitch UPD_LDA) and read by the four waves,
// the next tile's block being loaded into registers meanwhile, so A is read
// from L2 once per 128 columns of C; each wave keeps its 64 x 32 block of B
// in registers for the whole sweep.  A wave takes its row tile as four 16-row
// sub-blocks and prefetches the next sub-block's C while the current one is
// multiplied.  MFMA q of a sub-block uses k = 16 (lane / 16) + q, so a lane's
// A fragment is eight ds_read_b128 of one LDS row.  The MFMA computes the
// transposed block: lane l holds C(r + l%16, c + 4v + l/16), i.e. each C
// load / store instruction moves four 128-byte column runs.  Every global
// load is unconditional (clamped rows and columns) so that the vector-memory
// waits count only loads issued earlier.
// ---------------------------------------------------------------------------
#define UPD_LDA 66

// With FIRST the old C is read from the caller's matrix (fs: the permuted
// column list, rows >= fs.mtrue as zero) and the new C is written to C: the
// FP64 loop's first update, which fills the working copy.
template <bool NORMS, bool FIRST>
__global__ __launch_bounds__(256, 2) void k_gemm_upd64(int64_t M, DimSpec dN, DimSpec dK, int T,
                                                       int ncol, int nrg,
                                                       const double* __restrict__ A0, int64_t lda,
                                                       int64_t acoff, const int* acoff_dyn,
                                                       const double* __restrict__ Bt, int64_t ldbt,
                                                       double* __restrict__ C0, int64_t ldc,
                                                       int64_t ccoff, const int* ccoff_dyn,
                                                       double* __restrict__ normpart, int64_t ldnp,
                                                       FirstSrc fs) {
  __shared__ double As[2][64 * UPD_LDA];
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
  const double* A = A0 + (acoff + (acoff_dyn ? *acoff_dyn : 0)) * lda;
  // this wave's columns cw .. cw + ncw - 1; a wave past N keeps the workgroup's
  // first column as a (never stored) load target
  const int64_t cw = c0 + 32 * wid;
  const int ncw = (int)max((int64_t)0, min((int64_t)32, N - cw));
  double* C = C0 + (ccoff + (ccoff_dyn ? *ccoff_dyn : 0) + (ncw > 0 ? cw : c0)) * ldc;
  const int cmax = max(ncw, 1) - 1;
  // FIRST: the source column of each of the lane's eight C columns
  const double* cin[2][4];
  if (FIRST) {
    const int* fcols = fs.cols + fs.coff + (fs.coff_dyn ? (int64_t)*fs.coff_dyn : 0) + (ncw > 0 ? cw : c0);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) cin[h][v] = fs.p + (int64_t)fcols[min(16 * h + 4 * v + lk, cmax)] * fs.ld;
  }
  // B block of this wave: b0[q] = B(16 lk + q, cw + li), b1[q] = B(16 lk + q, cw + 16 + li)
  double b0[16], b1[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int64_t k = 16 * lk + q;
    b0[q] = (k < K && li < ncw) ? Bt[k * ldbt + cw + li] : 0.0;
    b1[q] = (k < K && 16 + li < ncw) ? Bt[k * ldbt + cw + 16 + li] : 0.0;
  }
  // A stager: column ka (clamped to K - 1: B is zero there), rows ar .. ar + 15 of a tile
  const int ka = tid >> 2, ar = 16 * (tid & 3);
  const double* acol = A + min((int64_t)ka, K - 1) * lda + ar;
  // the next tile's A goes through registers in two halves of eight rows
  // (loaded during sub-blocks 0 and 2, written to LDS during 1 and 3)
  double2 s0, s1, s2, s3;
#define UPD_LOADA(row0, half)                            \
  do {                                                   \
    const double* ap_ = acol + (row0) + 8 * (half);      \
    s0 = *(const double2*)(ap_);                         \
    s1 = *(const double2*)(ap_ + 2);                     \
    s2 = *(const double2*)(ap_ + 4);                     \
    s3 = *(const double2*)(ap_ + 6);                     \
  } while (0)
#define UPD_PUT(dst, j, v)                               \
  do {                                                   \
    (dst)[(ar + 2 * (j)) * UPD_LDA + ka] = (v).x;        \
    (dst)[(ar + 2 * (j) + 1) * UPD_LDA + ka] = (v).y;    \
  } while (0)
#define UPD_STOREA(bb, half)                             \
  do {                                                   \
    double* d_ = &As[bb][0] + 8 * (half) * UPD_LDA;      \
    UPD_PUT(d_, 0, s0);                                  \
    UPD_PUT(d_, 1, s1);                                  \
    UPD_PUT(d_, 2, s2);                                  \
    UPD_PUT(d_, 3, s3);                                  \
  } while (0)
  // C of one 16-row sub-block: cv[h][v] = C(r + li, 16 h + 4 v + lk)
#define UPD_LOADC(cv, r)                                                          \
  do {                                                                            \
    _Pragma("unroll") for (int h = 0; h < 2; ++h)                                 \
      _Pragma("unroll") for (int v = 0; v < 4; ++v) {                             \
        if (FIRST)                                                                \
          cv[h][v] = *first_src(cin[h][v], (r) + li, fs.mtrue);                   \
        else                                                                      \
          cv[h][v] = C[(r) + li + (int64_t)min(16 * h + 4 * v + lk, cmax) * ldc]; \
      }                                                                           \
  } while (0)

  double nsq[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int v = 0; v < 4; ++v) nsq[h][v] = 0.0;
  UPD_LOADA(rbase, 0);
  UPD_STOREA(0, 0);
  UPD_LOADA(rbase, 1);
  UPD_STOREA(0, 1);
  double cc[2][4], cn[2][4];
  UPD_LOADC(cc, rbase);
  __syncthreads();
  int buf = 0;
  for (int t = 0; t < T; ++t) {
    const int64_t row0 = rbase + (int64_t)t * 64;
    if (row0 >= M) break;  // uniform
    const bool more = t + 1 < T && row0 + 64 < M;
    const int64_t rnext = min(row0 + 64, M - 64);
    const double* as = &As[buf][0];
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
      const int64_t r = row0 + 16 * sb;
      if (sb == 0 || sb == 2) UPD_LOADA(rnext, sb >> 1);
      if ((sb == 1 || sb == 3) && more) UPD_STOREA(buf ^ 1, sb >> 1);
      UPD_LOADC(cn, sb < 3 ? r + 16 : rnext);
      const double* arow = as + (16 * sb + li) * UPD_LDA + 16 * lk;
      v4d acc0 = (v4d){0.0, 0.0, 0.0, 0.0}, acc1 = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 16; q += 2) {
        const double2 a = *(const double2*)(arow + q);
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(b0[q], a.x, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b1[q], a.x, acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(b0[q + 1], a.y, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b1[q + 1], a.y, acc1, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ca = 4 * v + lk, cb = 16 + 4 * v + lk;
        const double na = cc[0][v] - acc0[v], nb = cc[1][v] - acc1[v];
        if (ca < ncw) {
          C[r + li + (int64_t)ca * ldc] = na;
          if (NORMS) nsq[0][v] = fma(na, na, nsq[0][v]);
        }
        if (cb < ncw) {
          C[r + li + (int64_t)cb * ldc] = nb;
          if (NORMS) nsq[1][v] = fma(nb, nb, nsq[1][v]);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < 4; ++v) cc[h][v] = cn[h][v];
    }
    __syncthreads();
    buf ^= 1;
  }
#undef UPD_LOADA
#undef UPD_PUT
#undef UPD_STOREA
#undef UPD_LOADC
  if (NORMS) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double x = nsq[h][v];
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        x += __shfl_xor(x, 8, 64);
        const int cl = 16 * h + 4 * v + lk;
        if (li == 0 && cl < ncw) normpart[by * ldnp + cw + cl] = x;
      }
  }
}

bool rbrp_gemm_trail_first_ok(int64_t M, int64_t Kmax, int64_t lda) {
  return Kmax <= 64 && M % 64 == 0 && lda % 2 == 0;
}

hipError_t rbrp_launch_gemm_trail(hipStream_t st, int64_t M, DimSpec N, int64_t Nmax, DimSpec K,
                                  int64_t Kmax, const double* A, int64_t lda, int64_t acoff,
                                  const int* acoff_dyn, const double* Bt, int64_t ldbt, double* C,
                                  int64_t ldc, int64_t ccoff, const int* ccoff_dyn,
                                  double* normpart, int64_t ldnp, int64_t* groups, hipEvent_t e0,
                                  hipEvent_t e1, const FirstSrc* first) {
  if (groups) *groups = 0;
  if (M <= 0 || Nmax <= 0) return hipSuccess;
  if (Kmax > 128) return hipErrorInvalidValue;
  if (first && (!rbrp_gemm_trail_first_ok(M, Kmax, lda) || ((uintptr_t)A & 15) || !first->p ||
                !first->cols))
    return hipErrorInvalidValue;
  // (the 128-column kernel reads A in 16-byte units: A's columns must be
  // 16-byte aligned)
  if (Kmax <= 64 && M % 64 == 0 && lda % 2 == 0 && ((uintptr_t)A & 15) == 0) {
    const FirstSrc fs = first ? *first : FirstSrc{nullptr, 0, nullptr, 0, nullptr, 0};
    // 128-column workgroups with the A block shared through LDS
    const int64_t ncol = (Nmax + 127) / 128, nrow = M / 64;
    const int64_t T = std::max<int64_t>(
        1, std::min<int64_t>(RBRP_TRAIL_TMAX, (ncol * nrow) / (RBRP_TRAIL_WGS / 4)));
    const int64_t nrg = (nrow + T - 1) / T;
    dim3 grid((unsigned)rbrp_xcd_grid(ncol * nrg));
    if (groups) *groups = nrg;
    if (first && normpart)
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_upd64<true, true>), grid, dim3(256), 0, st, M, N, K, (int)T,
                     (int)ncol, (int)nrg, A, lda, acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff,
                     ccoff_dyn, normpart, ldnp, fs);
    else if (first)
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_upd64<false, true>), grid, dim3(256), 0, st, M, N, K, (int)T,
                     (int)ncol, (int)nrg, A, lda, acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff,
                     ccoff_dyn, normpart, ldnp, fs);
    else if (normpart)
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_upd64<true, false>), grid, dim3(256), 0, st, M, N, K, (int)T,
                     (int)ncol, (int)nrg, A, lda, acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff,
                     ccoff_dyn, normpart, ldnp, fs);
    else
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_upd64<false, false>), grid, dim3(256), 0, st, M, N, K, (int)T,
                     (int)ncol, (int)nrg, A, lda, acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff,
                     ccoff_dyn, normpart, ldnp, fs);
    return hipGetLastError();
  }
  const bool small = Kmax <= 64;
  const int64_t jb = small ? RBRP_TRAIL_JB : 1;  // rows per wave / 16 of the variant launched
  const int64_t ncol = (Nmax + 31) / 32, nrow = (M + 64 * jb - 1) / (64 * jb);
  int64_t T = std::max<int64_t>(
      1, std::min<int64_t>(RBRP_TRAIL_TMAX, (ncol * nrow) / RBRP_TRAIL_WGS));
  dim3 grid((unsigned)ncol, (unsigned)((nrow + T - 1) / T));
  if (groups) *groups = (int64_t)grid.y;
  if (small) {
    if (normpart)
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_trail<RBRP_TRAIL_JB, 16, true>), grid, dim3(256), 0, st, M, N, K, (int)T, A, lda,
                         acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff, ccoff_dyn, normpart, ldnp);
    else
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_trail<RBRP_TRAIL_JB, 16, false>), grid, dim3(256), 0, st, M, N, K, (int)T, A, lda,
                         acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff, ccoff_dyn, normpart, ldnp);
  } else {
    if (normpart)
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_trail<1, 32, true>), grid, dim3(256), 0, st, M, N, K, (int)T, A, lda,
                         acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff, ccoff_dyn, normpart, ldnp);
    else
      RBRP_LAUNCH_EV(e0, e1, (k_gemm_trail<1, 32, false>), grid, dim3(256), 0, st, M, N, K, (int)T, A, lda,
                         acoff, acoff_dyn, Bt, ldbt, C, ldc, ccoff, ccoff_dyn, normpart, ldnp);
  }
  return hipGetLastError();
}
