// Panel kernels: the pivoted Cholesky + robust filter of the sampled block's
// Gram matrix, and the second Cholesky-QR pass (with R2^{-1} and R2 * Rb).
// Blocks of at most 64 columns run on one wave (below); wider blocks (up to
// 128) on a 1024-thread workgroup with the Schur complement in registers
// (k_pivchol_regs, k_chol2_regs further down).
//
// Lane c owns column c.  The factor is kept transposed in LDS, one row per
// column (RT[c][t] = R(t, c), pitch PW_LDE).  Step j of the pivoted
// factorisation
//   R(j, c) = (G(P, c) - sum_{t<j} R(t, P) R(t, c)) / R(j, j)
// runs in groups of eight: the terms of the steps before the group come from
// the Schur complement G - R(0:j0, :)^T R(0:j0, :), kept in LDS and updated on
// the matrix core after each group (one LDS read per step), the group's own
// terms from the registers of the pivot's lane.  The second Cholesky-QR pass
// (no pivoting) reads the rows of RT for the earlier groups instead.  The outer
// loops stay rolled: a single wave runs these kernels, so their speed depends
// on the code fitting the instruction cache.  The pivot is the
// largest Schur diagonal among the columns not yet chosen, the lowest position
// on ties (a wave max and a ballot; the position tie-break runs only when
// values tie); the Schur diagonals are downdated as d -= R(j, c)^2 by a fused
// multiply-add.  The triangular inverses and R2 * Rb run as 16 x 16 block
// products on the FP64 matrix core.
//
// Reference: the pivoted QR of the sampled block and its filter
// (gpu/rbrp.cu:166-234, rbrp.hpp CPQR + filter), CholQR2 of the kept columns
// (gpu/rbrp.cu:236-266).
#include <stdint.h>

#include "rbrp_kernels.h"

#define PW_N 64
#define PW_LDO 65  // odd row pitch (doubles): lane-strided row access is conflict-free
#define PW_LDE 66  // even row pitch: rows are 16-byte aligned for double2 loads
#define PW_TILE_O (PW_N * PW_LDO)
#define PW_TILE_E (PW_N * PW_LDE)

namespace {

template <int CTRL>
__device__ __forceinline__ int pw_dpp(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double pw_dpp(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = pw_dpp<CTRL>((int)b), hi = pw_dpp<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Lanes l and l ^ 16 (W = 16) or l ^ 32 (W = 32): each lane receives its own
// value in one slot and its partner's in the other.
template <int W>
__device__ __forceinline__ double pw_swap_max(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)b, hi = (int)(b >> 32);
  auto sl = W == 16 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                    : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  auto sh = W == 16 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                    : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const double a = __longlong_as_double(((long long)sh[0] << 32) | (unsigned int)sl[0]);
  const double c = __longlong_as_double(((long long)sh[1] << 32) | (unsigned int)sl[1]);
  return fmax(v, fmax(a, c));
}
template <int W>
__device__ __forceinline__ int pw_swap_min(int v) {
  auto s = W == 16 ? __builtin_amdgcn_permlane16_swap(v, v, false, false)
                   : __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return min(v, min((int)s[0], (int)s[1]));
}
// Wave-wide maximum / minimum; every lane ends with the result.
__device__ __forceinline__ double pw_wave_max(double v) {
  v = fmax(v, pw_dpp<0x128>(v));  // row_ror:8
  v = fmax(v, pw_dpp<0x124>(v));  // row_ror:4
  v = fmax(v, pw_dpp<0x122>(v));  // row_ror:2
  v = fmax(v, pw_dpp<0x121>(v));  // row_ror:1
  v = pw_swap_max<16>(v);
  return pw_swap_max<32>(v);
}
__device__ __forceinline__ int pw_wave_min(int v) {
  v = min(v, pw_dpp<0x128>(v));
  v = min(v, pw_dpp<0x124>(v));
  v = min(v, pw_dpp<0x122>(v));
  v = min(v, pw_dpp<0x121>(v));
  v = pw_swap_min<16>(v);
  return pw_swap_min<32>(v);
}
__device__ __forceinline__ double pw_readlane(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lane c zeroes its row of an even-pitch tile (PW_LDE doubles).
__device__ __forceinline__ void pw_zero_row(double* T) {
  double2* r = reinterpret_cast<double2*>(T + (threadIdx.x & 63) * PW_LDE);
#pragma unroll 11
  for (int q = 0; q < PW_LDE / 2; ++q) r[q] = make_double2(0.0, 0.0);
}
// Lane r loads row r of the column-major X (n x n, coalesced across lanes),
// every load in flight at once: v[q] = X(r, q), indices clamped to n - 1.
__device__ __forceinline__ void pw_fetch_row(const double* __restrict__ X, int64_t ldx, int n,
                                             double (&v)[PW_N]) {
  const int rc = min((int)(threadIdx.x & 63), n - 1);
#pragma unroll
  for (int q = 0; q < PW_N; ++q) v[q] = X[rc + (int64_t)min(q, n - 1) * ldx];
}
// Y(r, q) = T[r * PW_LDO + q] for r, q < n (zero below the diagonal when
// upper is set); lane r writes row r, wave w of nw the columns in groups of
// eight from 8 w on, every 8 nw.
__device__ __forceinline__ void pw_store_rows(const double* T, int n, bool upper,
                                              double* __restrict__ Y, int64_t ldy, int w = 0,
                                              int nw = 1) {
  const int r = threadIdx.x & 63;
  if (r >= n) return;
  double* y = Y + r;
  for (int q0 = 8 * w; q0 < n; q0 += 8 * nw) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = T[r * PW_LDO + q0 + u];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q0 + u < n) y[(int64_t)(q0 + u) * ldy] = (upper && r > q0 + u) ? 0.0 : v[u];
  }
}

// 1 / sqrt(x) for a positive normal x: the hardware estimate refined by two
// Newton steps (no division).
__device__ __forceinline__ double pw_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = fma(y, fma(-h * y, y, 0.5), y);
  y = fma(y, fma(-h * y, y, 0.5), y);
  return y;
}

// acc[q] += sum_{l in [l0, l1)} U[(r0 + q) * PW_LDE + l] * x[l] for q < 8: eight
// broadcast rows of U against one row x (each 16-byte load of x feeds eight
// rows; sixteen independent accumulation chains, even and odd l).  l0 is a
// multiple of 8; terms in [l1, round8(l1)) must vanish and round8(l1) <= PW_N.
__device__ __forceinline__ void pw_ldot8(const double* U, int r0, const double* x, int l0, int l1,
                                         double (&acc)[8]) {
  double e[8], o[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    e[q] = acc[q];
    o[q] = 0.0;
  }
  const double* u = U + r0 * PW_LDE;
#pragma unroll 4
  for (int l = l0; l < l1; l += 2) {
    const double2 xv = *reinterpret_cast<const double2*>(x + l);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double2 uv = *reinterpret_cast<const double2*>(u + q * PW_LDE + l);
      e[q] = fma(uv.x, xv.x, e[q]);
      o[q] = fma(uv.y, xv.y, o[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = e[q] + o[q];
}

// ---------------------------------------------------------------------------
// 16 x 16 block products on the FP64 matrix core for the one-wave kernels.
// v_mfma_f64_16x16x4f64 operands: A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k; the result D[i][j] sits in lane j + 16 (i % 4), element i / 4.  A
// 16-deep product is four MFMAs, step s taking k = 4 s + (lane / 16), so lane
// (j, g) = j + 16 g holds D[4 v + g][j] in element v, which is exactly its B
// operand of step v for a following product: a chain of products stays in
// registers.
// ---------------------------------------------------------------------------
// acc += A(I, K) * B(K, J) for one 16 x 16 block pair: A's row r at
// A + r * lda (row-major), B's column c at Bt + c * ldb (B transposed), block
// origins (ra, ka) and (kb, cb).
__device__ __forceinline__ v4d pw_mfma_block(v4d acc, const double* A, int lda, int ra, int ka,
                                             const double* Bt, int ldb, int kb, int cb) {
  const int j = threadIdx.x & 15, g = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double a = A[(ra + j) * lda + ka + 4 * s + g];
    const double b = Bt[(cb + j) * ldb + kb + 4 * s + g];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  return acc;
}

// X = U^{-1} for the n x n upper-triangular U (row r at U + r * PW_LDE, zero
// outside its n x n block; dinv[r] = 1 / U(r, r) for r < n), written to
// Xo[r * PW_LDO + c] for the blocks on and above the diagonal of the
// ceil(n / 16) x ceil(n / 16) block grid (the rows and columns in
// [n, 16 ceil(n / 16)) hold the identity; blocks below the diagonal are not
// written).  The diagonal 16 x 16 blocks are inverted by back substitution,
// lane group g (lanes 16 g .. 16 g + 15) on block g with lane j computing its
// column j; then, by distance d = J - I from the diagonal,
//   X(I, J) = -X(I, I) * sum_{K = I+1}^{J} U(I, K) X(K, J)
// on the matrix core, the sum's MFMA output feeding the second product as
// its B operand.  Called by every wave of the workgroup: wave 0 inverts the
// diagonal blocks, the blocks at one distance are spread over the nw waves (a
// barrier between distances), and Xo is complete when it returns.
__device__ void pw_upper_inverse_mfma(const double* U, const double* dinv, double* Xo, int n,
                                      int w = 0, int nw = 1) {
  const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
  const int nbk = (n + 15) >> 4;  // uniform
  if (w == 0 && g < nbk) {
    const int R0 = 16 * g;
    double x[16];
#pragma unroll
    for (int r = 15; r >= 0; --r) {
      const int R = R0 + r;
      double a = (r == j) ? 1.0 : 0.0;
#pragma unroll
      for (int p = r + 1; p < 16; ++p) a = fma(-U[R * PW_LDE + R0 + p], x[p], a);
      x[r] = (R < n) ? ((r <= j) ? a * dinv[R] : 0.0) : ((r == j) ? 1.0 : 0.0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) Xo[(R0 + r) * PW_LDO + R0 + j] = x[r];
  }
  for (int d = 1; d < nbk; ++d) {
    __syncthreads();  // the blocks nearer the diagonal are written
    for (int I = w; I + d < nbk; I += nw) {
      const int J = I + d;
      v4d y = (v4d){0.0, 0.0, 0.0, 0.0};
      for (int K = I + 1; K <= J; ++K) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const double a = U[(16 * I + j) * PW_LDE + 16 * K + 4 * s + g];
          const double b = Xo[(16 * K + 4 * s + g) * PW_LDO + 16 * J + j];
          y = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, y, 0, 0, 0);
        }
      }
      v4d xo = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double a = Xo[(16 * I + j) * PW_LDO + 16 * I + 4 * s + g];
        xo = __builtin_amdgcn_mfma_f64_16x16x4f64(a, y[s], xo, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) Xo[(16 * I + 4 * v + g) * PW_LDO + 16 * J + j] = -xo[v];
    }
  }
  __syncthreads();
}

// S -= Rg^T Rg for the eight factor rows t = t0 .. t0+7 (Rg(t, p) = RT[p *
// PW_LDE + t]) on the blocks (I, J), I <= J < nbk, of the 16 x 16 block grid of
// the row-major S (odd pitch), the diagonal blocks whole.  Lane (i, k) = i + 16 k
// holds x[I][h] = Rg(t0 + 4h + k, 16 I + i): the A operand of block row I and
// the B operand of block column I; the block's result sits in lane (j, g) as
// S(16 I + 4 v + g, 16 J + j), element v.
__device__ void pw_schur8(double* S, const double* RT, int t0, int nbk) {
  const int j = threadIdx.x & 15, g = (threadIdx.x & 63) >> 4;
  double x[4][2];
#pragma unroll
  for (int I = 0; I < 4; ++I)
#pragma unroll
    for (int h = 0; h < 2; ++h) x[I][h] = RT[(16 * I + j) * PW_LDE + t0 + 4 * h + g];
#pragma unroll
  for (int I = 0; I < 4; ++I)
#pragma unroll
    for (int J = I; J < 4; ++J) {
      if (J >= nbk) continue;  // uniform
      double* blk = S + (16 * I + g) * PW_LDO + 16 * J + j;
      v4d acc;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] = blk[4 * v * PW_LDO];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-x[I][0], x[J][0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-x[I][1], x[J][1], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) blk[4 * v * PW_LDO] = acc[v];
    }
}

}  // namespace

// LDS (doubles): Gs (odd pitch) | RT (even) | Ws (even) | Bt (even) | three
// vectors of PW_N.
#define PW_LDS_DOUBLES (PW_TILE_O + 3 * PW_TILE_E + 3 * PW_N)
// Waves of the two panel kernels below: the factorisation runs on the first;
// the stores of the 64 x 64 results, the triangular inverse and R2 * Rb are
// shared by all of them.
#define PW_NW 4

__global__ __launch_bounds__(64 * PW_NW) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_pivchol_wave(const double* __restrict__ G, int64_t ldg,
                                                     int b, int filter_mode, int64_t i,
                                                     const int* __restrict__ J,
                                                     RbrpBlockState* bs, double* __restrict__ Rb,
                                                     double* __restrict__ Rinv,
                                                     int* __restrict__ Jk, int* __restrict__ Jin,
                                                     int* __restrict__ Jout,
                                                     int* __restrict__ Jkpos) {
  extern __shared__ __align__(16) double lds[];
  double* Gs = lds;                // Gram rows, later store staging (odd pitch)
  double* RT = Gs + PW_TILE_O;     // RT[c][t] = R(t, c); later the inverse, transposed
  double* Ws = RT + PW_TILE_E;     // kept factor, row-major
  double* sqr = Ws + 2 * PW_TILE_E;
  double* dinv = sqr + PW_N;
  int* pos2col = reinterpret_cast<int*>(dinv + PW_N);
  __shared__ int kept[2];  // bt and the degenerate flag, from the first wave
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int s = bs->s;
  if (s <= 0) {
    if (threadIdx.x == 0) {
      bs->bt = 0;
      bs->nswap = 0;
    }
    return;
  }
  const int jv = J[min(c, s - 1)];  // sampled column index of lane c
  if (w == 0) {
    {
      double v[PW_N];
      pw_fetch_row(G, ldg, s, v);
      pw_zero_row(RT);
#pragma unroll
      for (int q = 0; q < PW_N; ++q) Gs[c * PW_LDO + q] = v[q];  // Gs[r][q] = G(r, q), r, q < s
    }
    const bool act = c < s;
    double d = act ? Gs[c * PW_LDO + c] : 0.0;
    int pos = c;
    int rank = s;
    double* myrow = RT + c * PW_LDE;
    // Left-looking in groups of eight steps.  Gs holds the Schur complement
    // S = G - R(0:j0, :)^T R(0:j0, :) of the steps before the group (blocks on and
    // above the diagonal of its 16 x 16 grid, updated on the matrix core after
    // each group); a step adds the group's earlier terms from the registers of
    // the pivot's lane.  Lane c reads S(P, c) from row P when P's block is at
    // most c's, else from column P (S is symmetric).
    const int nbk = (s + 15) >> 4;
    const int bc = c >> 4;
    for (int j0 = 0; j0 < s; j0 += 8) {
      double rv[8];  // rv[q] = R(j0 + q, c)
#pragma unroll
      for (int q = 0; q < 8; ++q) rv[q] = 0.0;
      bool stop = false;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int j = j0 + q;
        if (stop || j >= s) continue;  // uniform
        const double key = (act && pos >= j) ? d : -1.0;
        const double m = pw_wave_max(key);
        if (!(m > 0.0)) {
          rank = j;
          stop = true;
          continue;
        }
        const uint64_t ties = __builtin_amdgcn_ballot_w64(key == m);
        int P, bp;
        if (__builtin_popcountll(ties) == 1) {
          P = (int)__builtin_ctzll(ties);
          bp = __builtin_amdgcn_readlane(pos, P);
        } else {
          bp = pw_wave_min(key == m ? pos : 0x7fffffff);
          P = (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(key == m && pos == bp));
        }
        const double rjj = sqrt(m);
        const double inv = pw_rsqrt(m);
        const double g = Gs[((P >> 4) <= bc) ? P * PW_LDO + c : c * PW_LDO + P];
        double acc = 0.0;
#pragma unroll
        for (int p = 0; p < q; ++p) acc = fma(pw_readlane(rv[p], P), rv[p], acc);
        // the column at position j moves to the pivot's position, the pivot to j
        if (pos == j) pos = bp;
        if (c == P) pos = j;
        double val = 0.0;
        if (act) {
          if (c == P) {
            val = rjj;
          } else if (pos > j) {
            val = (g - acc) * inv;
            d = fma(-val, val, d);
          }
        }
        rv[q] = val;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) myrow[j0 + q] = rv[q];
      if (stop || j0 + 8 >= s) break;  // uniform
      pw_schur8(Gs, RT, j0, nbk);
    }
    if (act) pos2col[pos] = c;
    // energy of factor row r (rows at and beyond the rank are zero)
    {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      if (c < rank) {
#pragma unroll 4
        for (int q = 0; q < PW_N; q += 4) {
          const double a0 = RT[q * PW_LDE + c], a1 = RT[(q + 1) * PW_LDE + c];
          const double a2 = RT[(q + 2) * PW_LDE + c], a3 = RT[(q + 3) * PW_LDE + c];
          s0 = fma(a0, a0, s0);
          s1 = fma(a1, a1, s1);
          s2 = fma(a2, a2, s2);
          s3 = fma(a3, a3, s3);
        }
      }
      sqr[c] = (s0 + s1) + (s2 + s3);
    }
    // robust filter: csum over the rows from the last one up (the same serial
    // order in every lane); keep bt = s - k for the first k with csum >= thr
    double total = 0.0;
    for (int q0 = s - 1; q0 >= 0; q0 -= 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = sqr[max(q0 - u, 0)];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 - u >= 0) total += v[u];
    }
    const double thr = total / (double)(filter_mode == 0 ? s : b);
    int k = s;
    {
      double csum = 0.0;
      for (int q0 = 0; q0 < s && k == s; q0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = sqr[max(s - 1 - q0 - u, 0)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (k == s && q0 + u < s) {
            csum += v[u];
            if (csum >= thr) k = q0 + u;
          }
        }
      }
    }
    const bool degen = rank == 0 || !(total > 0.0);
    int bt0 = degen ? 0 : s - k;
    if (bt0 > rank) bt0 = rank;
    if (c == 0) {
      kept[0] = bt0;
      kept[1] = degen ? 1 : 0;
    }
  }  // first wave
  __syncthreads();
  const int bt = kept[0];
  const bool degenerate = kept[1] != 0;
  // kept factor in pivot order: column p = column pos2col[p] of R, rows <= p,
  // row-major into Ws (broadcast rows) and Gs (store staging), rows 8 w .. by
  // wave
  const int col = pos2col[min(c, max(bt - 1, 0))];
  {
    const double* src = RT + col * PW_LDE;
    for (int t0 = 8 * w; t0 < PW_N; t0 += 8 * PW_NW) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = src[t0 + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double x = (c < bt && t0 + u <= c) ? v[u] : 0.0;
        Ws[(t0 + u) * PW_LDE + c] = x;
        Gs[(t0 + u) * PW_LDO + c] = x;
      }
    }
    if (w == 0 && c < bt) dinv[c] = 1.0 / src[c];
  }
  __syncthreads();
  pw_store_rows(Gs, bt, true, Rb, b, w, PW_NW);
  __syncthreads();
  pw_upper_inverse_mfma(Ws, dinv, Gs, bt, w, PW_NW);
  pw_store_rows(Gs, bt, true, Rinv, b, w, PW_NW);
  if (w != PW_NW - 1) return;  // (the last wave has no inverse blocks of its own)
  // index sets: Jk (pivot order), Jin = sort(Jk) beyond the block, Jout =
  // block positions not in Jk (ascending)
  int v = 0x7fffffff;
  {
    const int jc = __builtin_amdgcn_ds_bpermute(col << 2, jv);  // J[col] from lane col
    if (c < bt) {
      v = jc;
      Jk[c] = v;
      if (Jkpos) Jkpos[c] = col;
    }
  }
  const int q = (int)(i + c);
  int rk = 0, member = 0;
  for (int a = 0; a < bt; ++a) {
    const int va = __builtin_amdgcn_readlane(v, a);
    rk += (va < v) ? 1 : 0;
    member |= (va == q) ? 1 : 0;
  }
  const bool flag = c < bt && !member;
  const int nin = (int)__builtin_popcountll(
      __builtin_amdgcn_ballot_w64(c < bt && (int64_t)v >= i + bt));
  if (c < bt && rk >= bt - nin) Jin[rk - (bt - nin)] = v;
  const uint64_t fl = __builtin_amdgcn_ballot_w64(flag);
  if (flag) Jout[__builtin_popcountll(fl & ((1ull << c) - 1ull))] = q;
  if (c == 0) {
    const int nout = (int)__builtin_popcountll(fl);
    if (degenerate) bs->status = 1;
    if (nin != nout) bs->status = 3;
    bs->bt = bt;
    bs->nswap = nin;
  }
}

// Second Cholesky-QR pass fused with the new diagonal block of R: G2 = R2^T R2,
// R2inv = R2^{-1}, Rout = R2 * Rb (bt x bt, upper triangular, zeros below).
// The first wave factorises while the second stages Rb; the products, the
// inverse and the stores are spread over the four waves.
__global__ __launch_bounds__(64 * PW_NW) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_chol2_wave(const double* __restrict__ G2, int64_t ldg,
                                                   int b, RbrpBlockState* bs,
                                                   const double* __restrict__ Rb,
                                                   double* __restrict__ R2,
                                                   double* __restrict__ R2inv,
                                                   double* __restrict__ Rout, int64_t ldr) {
  extern __shared__ __align__(16) double lds[];
  double* Gs = lds;               // G2 rows, later store staging (odd pitch)
  double* RT = Gs + PW_TILE_O;    // RT[c][t] = R2(t, c); later the inverse, transposed
  double* Ws = RT + PW_TILE_E;    // R2, row-major
  double* Bt = Ws + PW_TILE_E;    // Bt[c][l] = Rb(l, c)
  double* dinv = Bt + PW_TILE_E;
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int bt = bs->bt;
  if (bt <= 0) return;
  if (w == 1) {
    // Rb transposed: Bt[q][c] = Rb(c, q) (zero below the diagonal and beyond bt)
    double v[PW_N];
    pw_fetch_row(Rb, b, bt, v);
#pragma unroll
    for (int q = 0; q < PW_N; ++q)
      Bt[q * PW_LDE + c] = (c < bt && q < bt && c <= q) ? v[q] : 0.0;
  }
  if (w == 0) {
    {
      double v[PW_N];
      pw_fetch_row(G2, ldg, bt, v);
      pw_zero_row(RT);
#pragma unroll
      for (int q = 0; q < PW_N; ++q) Gs[c * PW_LDO + q] = v[q];  // G2(r, q), r, q < bt
    }
    double d = c < bt ? Gs[c * PW_LDO + c] : 0.0;
    bool bad = false;
    double* myrow = RT + c * PW_LDE;
    // left-looking by blocks of eight columns: the terms of earlier blocks for
    // all eight at once, then the in-block terms from the registers of lane j
    for (int j0 = 0; j0 < bt; j0 += 8) {
      double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      pw_ldot8(RT, j0, myrow, 0, j0, acc);  // RT row j = column j of R2
      double g[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) g[q] = Gs[(j0 + q) * PW_LDO + c];
      double rv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int j = j0 + q;
        rv[q] = 0.0;
        if (j < bt) {
          const double dj = pw_readlane(d, j);
          double a = acc[q];
#pragma unroll
          for (int p = 0; p < q; ++p) a = fma(pw_readlane(rv[p], j), rv[p], a);
          const double dd = dj > 0.0 ? dj : 1.0;
          const double rjj = sqrt(dd);
          const double inv = pw_rsqrt(dd);
          bad |= !(dj > 0.0);
          if (c < bt) {
            if (c == j) {
              rv[q] = rjj;
            } else if (c > j) {
              rv[q] = (g[q] - a) * inv;
              d = fma(-rv[q], rv[q], d);
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) myrow[j0 + q] = rv[q];
    }
    if (c == 0 && bad) bs->status = 2;
    if (c < bt) dinv[c] = 1.0 / myrow[c];
  }
  __syncthreads();
  // R2 row-major into Ws (broadcast rows) and Gs (store staging), rows 8 w ..
  // by wave
  for (int t0 = 8 * w; t0 < PW_N; t0 += 8 * PW_NW) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = RT[c * PW_LDE + t0 + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double x = (c < bt && t0 + u < bt) ? v[u] : 0.0;
      Ws[(t0 + u) * PW_LDE + c] = x;
      Gs[(t0 + u) * PW_LDO + c] = x;
    }
  }
  __syncthreads();
  pw_store_rows(Gs, bt, true, R2, b, w, PW_NW);
  __syncthreads();
  // Rout = R2 * Rb on the matrix core, block (I, J) = sum_{K=I}^{J} R2(I, K)
  // Rb(K, J): the ten blocks of the 4 x 4 grid spread over the waves by their
  // J - I + 1 products (five per wave)
  {
    const int nbk = (bt + 15) >> 4, j = c & 15, g = c >> 4;
    constexpr int kI[10] = {0, 0, 0, 0, 1, 1, 2, 1, 2, 3};
    constexpr int kJ[10] = {3, 0, 2, 1, 3, 2, 3, 1, 2, 3};
    constexpr int kW[10] = {0, 0, 1, 1, 2, 2, 3, 3, 3, 3};
#pragma unroll
    for (int p = 0; p < 10; ++p) {
      const int I = kI[p], J = kJ[p];
      if (kW[p] != w || J >= nbk) continue;  // uniform
      v4d acc = (v4d){0.0, 0.0, 0.0, 0.0};
      for (int K = I; K <= J; ++K)
        acc = pw_mfma_block(acc, Ws, PW_LDE, 16 * I, 16 * K, Bt, PW_LDE, 16 * K, 16 * J);
#pragma unroll
      for (int v = 0; v < 4; ++v) Gs[(16 * I + 4 * v + g) * PW_LDO + 16 * J + j] = acc[v];
    }
  }
  __syncthreads();
  pw_store_rows(Gs, bt, true, Rout, ldr, w, PW_NW);
  __syncthreads();
  pw_upper_inverse_mfma(Ws, dinv, Gs, bt, w, PW_NW);
  pw_store_rows(Gs, bt, true, R2inv, b, w, PW_NW);
}

// Inverses of the diagonal blocks of a k x k upper-triangular R, in the
// partition of the blocked back substitution: super-blocks of 256 rows from the
// bottom (s1 = k - 256 j, s0 = max(0, s1 - 256)), each cut into 64-row blocks
// from its bottom (r1 = s1 - 64 i, r0 = max(s0, r1 - 64)).  Workgroup 4 j + i
// writes the inverse of R(r0:r1, r0:r1) to out + (4 j + i) * 4096 (leading
// dimension 64, zeros below the diagonal); blocks that do not exist are
// skipped.  One wave per block, as in the panel kernels.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_upper_inv_blocks(
    int64_t k, const double* __restrict__ R, int64_t ldr, double* __restrict__ out) {
  extern __shared__ __align__(16) double lds[];
  double* Gs = lds;                // inverse rows (store staging, odd pitch)
  double* RT = Gs + PW_TILE_O;     // inverse, transposed (zero on entry)
  double* Ws = RT + PW_TILE_E;     // the block, row-major
  double* dinv = Ws + 2 * PW_TILE_E;
  const int64_t j = blockIdx.x >> 2, ib = blockIdx.x & 3;
  const int64_t s1 = k - 256 * j, s0 = max((int64_t)0, s1 - 256);
  const int64_t r1 = s1 - 64 * ib;
  if (r1 <= s0) return;  // uniform
  const int64_t r0 = max(s0, r1 - 64);
  const int nb = (int)(r1 - r0);
  const int c = threadIdx.x;
  const double* rc = R + r0 + (r0 + min(c, nb - 1)) * ldr;  // column c of the block
  for (int t0 = 0; t0 < PW_N; t0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = rc[min(t0 + u, nb - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      Ws[(t0 + u) * PW_LDE + c] = (c < nb && t0 + u <= c) ? v[u] : 0.0;
  }
  if (c < nb) dinv[c] = 1.0 / Ws[c * PW_LDE + c];
  pw_upper_inverse_mfma(Ws, dinv, Gs, nb);
  pw_store_rows(Gs, nb, true, out + (int64_t)blockIdx.x * 4096, 64);
}

// ---------------------------------------------------------------------------
// Workgroup panel kernels for blocks of up to SMAX = 128 columns.
//
// The Schur complement of the s x s Gram matrix lives in registers: thread
// (a, t) of the (SMAX / 4)^2 threads holds the 4 x 4 tile of rows 4a .. 4a+3
// and columns 4t .. 4t+3.  Rows and columns keep their sampled order; the
// dgeqp3 positions are a permutation only.  Step j of the right-looking
// pivoted factorisation (l = S(P, :) / r_jj, S -= l l^T):
//   1. every wave finds the pivot from its own copy of the Schur diagonal
//      (lane q holds entries q and q + 64 with their positions): the largest
//      value among the columns not yet chosen, the lowest position on ties;
//   2. the threads holding row P write l (zero in the chosen columns and at P)
//      to one of two LDS vectors, and row j of the factor (sampled column
//      order) to LDS;
//   3. one barrier; every thread updates its tile, every wave its copy of the
//      diagonal (d -= l .* l).
// Alternating the two l vectors makes one barrier per step enough: the vector
// written at step j + 1 was last read before the barrier of step j.  The
// arithmetic per entry (a division by r_jj, one fused multiply-add per step)
// is that of the textbook outer-product Cholesky.
// ---------------------------------------------------------------------------
template <int SMAX>
struct PrShape {
  static constexpr int TPR = SMAX / 4;  // column tiles per row band
  static constexpr int NT = TPR * TPR;  // threads
  static constexpr int H = SMAX / 64;   // diagonal entries per lane
};

// The thread's 4 x 4 tile, element (q, e) at 4 q + e.  A vector value rather
// than an array: an element picked by a uniform runtime index is read with
// register-indexed moves, without branches and without leaving registers.
typedef double v16d __attribute__((ext_vector_type(16)));

// f(r, c) for every r, c < n (n <= SMAX): indices from the SMAX x n block by
// shifts and masks (no division by n), r fastest (coalesced column-major
// stores).
template <int SMAX, typename Fn>
__device__ __forceinline__ void pr_for_square(int n, Fn f) {
  for (int idx = threadIdx.x; idx < SMAX * n; idx += blockDim.x) {
    const int r = idx % SMAX, c = idx / SMAX;
    if (r < n) f(r, c);
  }
}

// The n x n column-major X (leading dimension ldx) into the register tiles,
// staged through F (row r at F + r * ld) for coalesced loads; zero outside
// n x n.  F is free again after the caller's next barrier.
template <int SMAX>
__device__ __forceinline__ v16d pr_load_tiles(const double* __restrict__ X, int64_t ldx, int n,
                                              double* F, int ld) {
  constexpr int NT = PrShape<SMAX>::NT, TPR = PrShape<SMAX>::TPR, E = SMAX * SMAX / NT;
  const int tid = threadIdx.x, a = tid / TPR, t = tid % TPR;
  double v[E];  // every load in flight at once
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int idx = tid + k * NT, r = idx % SMAX, c = idx / SMAX;
    v[k] = X[min(r, n - 1) + (int64_t)min(c, n - 1) * ldx];  // clamped: no branches
  }
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int idx = tid + k * NT, r = idx % SMAX, c = idx / SMAX;
    if (r < n && c < n) F[r * ld + c] = v[k];
  }
  __syncthreads();
  v16d T;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = 4 * a + q, c = 4 * t + e;
      T[4 * q + e] = (r < n && c < n) ? F[r * ld + c] : 0.0;
    }
  return T;
}

// Wave-wide maximum of an unsigned value; every lane ends with the result.
template <int W>
__device__ __forceinline__ unsigned pr_swap_maxu(unsigned v) {
  auto x = W == 16 ? __builtin_amdgcn_permlane16_swap(v, v, false, false)
                   : __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return max(v, max((unsigned)x[0], (unsigned)x[1]));
}
__device__ __forceinline__ unsigned pr_wave_maxu(unsigned v) {
  v = max(v, (unsigned)pw_dpp<0x128>((int)v));  // row_ror:8
  v = max(v, (unsigned)pw_dpp<0x124>((int)v));  // row_ror:4
  v = max(v, (unsigned)pw_dpp<0x122>((int)v));  // row_ror:2
  v = max(v, (unsigned)pw_dpp<0x121>((int)v));  // row_ror:1
  v = pr_swap_maxu<16>(v);
  return pr_swap_maxu<32>(v);
}

// T -= l(rows) l(cols)^T for the thread's tile.
template <int SMAX>
__device__ __forceinline__ v16d pr_rank1(const double* l, v16d T) {
  constexpr int TPR = PrShape<SMAX>::TPR;
  const int a = threadIdx.x / TPR, t = threadIdx.x % TPR;
  double lr[4], lc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) lr[q] = l[4 * a + q];
#pragma unroll
  for (int e = 0; e < 4; ++e) lc[e] = l[4 * t + e];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) T[4 * q + e] = T[4 * q + e] - lr[q] * lc[e];
  return T;
}

// X(r, c) of the inverse being built by pr_upper_inverse: X(r, c) for r < c
// in F[c * ld + r], the diagonal in dinv, zero below the diagonal and outside
// n x n.
__device__ __forceinline__ double pr_xload(const double* F, int ld, const double* dinv, int n,
                                           int r, int c) {
  if (r >= n || c >= n || r > c) return 0.0;
  return r < c ? F[c * ld + r] : dinv[r];
}
// F(r, c), zero outside n x n.
__device__ __forceinline__ double pr_uload(const double* F, int ld, int n, int r, int c) {
  return (r < n && c < n) ? F[r * ld + c] : 0.0;
}

// X = U^{-1} for the n x n upper-triangular U in the upper triangle of F (row
// r at F + r * ld), n <= 128, written to Xo (column-major, ldx) with zeros
// below the diagonal.  X(l, c), l < c, is built in F[c * ld + l] (the strictly
// lower triangle, free once U is final) and 1 / U(c, c) in dinv.  The 16 x 16
// diagonal blocks are inverted by back substitution, lane group g = tid / 16
// on block g with lane j computing its column j; then, by distance d = J - I
// from the diagonal, one wave per block,
//   X(I, J) = -X(I, I) * sum_{K = I+1}^{J} U(I, K) X(K, J)
// on the FP64 matrix core (v_mfma_f64_16x16x4f64: A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[4 v + g][j] in lane j + 16 g, element v, so the
// sum's output is the second product's B operand as it stands).
template <int SMAX>
__device__ __forceinline__ void pr_upper_inverse(double* F, int ld, int n, double* dinv,
                                                 double* __restrict__ Xo, int64_t ldx) {
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int c = tid; c < n; c += NT) dinv[c] = 1.0 / F[c * ld + c];
  __syncthreads();
  const int nbk = (n + 15) >> 4;
  const int lane = tid & 63, j = lane & 15, g = lane >> 4, nw = NT >> 6;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, uniform
  if ((tid >> 4) < nbk) {
    const int R0 = 16 * (tid >> 4);
    double x[16];
#pragma unroll
    for (int r = 15; r >= 0; --r) {
      const int R = R0 + r;
      double acc = (r == j) ? 1.0 : 0.0;
#pragma unroll
      for (int p = r + 1; p < 16; ++p) {
        acc = fma(-pr_uload(F, ld, n, R, R0 + p), x[p], acc);
      }
      x[r] = (R < n && r <= j) ? acc * dinv[R] : 0.0;
    }
    if (R0 + j < n) {
#pragma unroll
      for (int r = 0; r < 16; ++r)
        if (r < j) F[(R0 + j) * ld + R0 + r] = x[r];
    }
  }
  __syncthreads();
  for (int d = 1; d < nbk; ++d) {
    for (int I = w; I + d < nbk; I += nw) {  // wave-uniform
      const int Jb = I + d;
      // four accumulators (one per k step of a 16-deep product): dependent
      // MFMAs wait out the whole pipeline, independent ones overlap
      v4d y4[4];
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) y4[s4] = (v4d){0.0, 0.0, 0.0, 0.0};
      for (int K = I + 1; K <= Jb; ++K) {
        double av[4], bv[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int ra = 16 * I + j, ca = 16 * K + 4 * s4 + g;
          av[s4] = pr_uload(F, ld, n, ra, ca);
          bv[s4] = pr_xload(F, ld, dinv, n, ca, 16 * Jb + j);
        }
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
          y4[s4] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s4], bv[s4], y4[s4], 0, 0, 0);
      }
      const v4d y = (y4[0] + y4[1]) + (y4[2] + y4[3]);
      v4d x4[4];
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const double av = pr_xload(F, ld, dinv, n, 16 * I + j, 16 * I + 4 * s4 + g);
        x4[s4] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, y[s4], (v4d){0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
      }
      const v4d xo = (x4[0] + x4[1]) + (x4[2] + x4[3]);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int xr = 16 * I + 4 * v + g, xc = 16 * Jb + j;
        if (xr < n && xc < n) F[xc * ld + xr] = -xo[v];
      }
    }
    __syncthreads();
  }
  pr_for_square<SMAX>(n, [&](int r, int c) {
    Xo[r + (int64_t)c * ldx] = (r < c) ? F[c * ld + r] : ((r == c) ? dinv[c] : 0.0);
  });
}

// Pivoted Cholesky of the sampled block's Gram matrix with the robust filter
// (the outputs of k_pivchol_wave, for s <= SMAX).
template <int SMAX>
__global__ __launch_bounds__(PrShape<SMAX>::NT) void k_pivchol_regs(
    const double* __restrict__ G, int64_t ldg, int b, int filter_mode, int64_t i,
    const int* __restrict__ J, RbrpBlockState* bs, double* __restrict__ Rb,
    double* __restrict__ Rinv, int* __restrict__ Jk, int* __restrict__ Jin,
    int* __restrict__ Jout, int* __restrict__ Jkpos) {
  constexpr int NT = PrShape<SMAX>::NT, TPR = PrShape<SMAX>::TPR, H = PrShape<SMAX>::H;
  extern __shared__ double F[];  // s x (s + 1): the Gram matrix, then the factor
  __shared__ double lv[2][SMAX];
  __shared__ double sqr[SMAX], dinv[SMAX], dpv[SMAX];
  __shared__ int piv[SMAX], sj[SMAX], sorted[SMAX], flag[SMAX];
  __shared__ int s_bt;
  const int tid = threadIdx.x, lane = tid & 63, a = tid / TPR, t = tid % TPR;
  const int s = bs->s;
  if (s <= 0) {
    if (tid == 0) {
      bs->bt = 0;
      bs->nswap = 0;
    }
    return;
  }
  const int ld = s + 1;
  v16d T = pr_load_tiles<SMAX>(G, ldg, s, F, ld);
  double dg[H];
  int ps[H];  // position of column lane + 64 h (-1 beyond s)
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const int c = lane + 64 * h;
    dg[h] = c < s ? F[c * ld + c] : 0.0;
    ps[h] = c < s ? c : -1;
  }
  __syncthreads();  // F receives factor rows from here on
  unsigned chosen = 0;  // bit e: column 4t + e is a pivot
  int rank = s;
  for (int j = 0; j < s; ++j) {
    // The pivot: the candidates' values are positive doubles, whose bit
    // patterns order like the values, so the wave maximum runs on the high
    // words and then on the low words of the lanes that hold the high maximum.
    uint64_t kb = 0;  // bits of the lane's best candidate (0: none)
    int lp = 0x7fffffff, lc = 0;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const uint64_t k = (ps[h] >= j && dg[h] > 0.0) ? (uint64_t)__double_as_longlong(dg[h]) : 0;
      if (k > kb || (k == kb && k != 0 && ps[h] < lp)) {
        kb = k;
        lp = ps[h];
        lc = lane + 64 * h;
      }
    }
    const unsigned hi = (unsigned)(kb >> 32), lo = (unsigned)kb;
    const unsigned mh = pr_wave_maxu(hi);
    const unsigned ml = pr_wave_maxu(hi == mh ? lo : 0u);
    if (mh == 0 && ml == 0) {  // no positive diagonal left (uniform: all waves agree)
      rank = j;
      break;
    }
    const bool top = hi == mh && lo == ml;
    const uint64_t ties = __builtin_amdgcn_ballot_w64(top);
    int P, bp;
    if (__builtin_popcountll(ties) == 1) {
      const int L = (int)__builtin_ctzll(ties);
      bp = __builtin_amdgcn_readlane(lp, L);
      P = __builtin_amdgcn_readlane(lc, L);
    } else {
      bp = pw_wave_min(top ? lp : 0x7fffffff);
      const int L = (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(top && lp == bp));
      P = __builtin_amdgcn_readlane(lc, L);
    }
    const double m = __longlong_as_double((long long)(((uint64_t)mh << 32) | ml));
    double* l = lv[j & 1];
    if (a == (P >> 2)) {
      const int q = P & 3;
      const double inv = pw_rsqrt(m);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = 4 * t + e;
        const bool zero = c >= s || ((chosen >> e) & 1u) || c == P;
        const double x = T[4 * q + e] * inv;
        const double v = zero ? 0.0 : x;
        l[c] = v;
        if (c < s) F[j * ld + c] = v;  // (the diagonal entry is set after the loop)
        if (c == P) dpv[j] = m;
      }
    }
    __syncthreads();
    T = pr_rank1<SMAX>(l, T);
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const double x = l[lane + 64 * h];
      dg[h] = dg[h] - x * x;
      // the column at position j moves to the pivot's position, the pivot to j
      if (ps[h] == j) ps[h] = bp;
      if (lane + 64 * h == P) ps[h] = j;
    }
    if ((P >> 2) == t) chosen |= 1u << (P & 3);
  }
  if (tid < 64) {
#pragma unroll
    for (int h = 0; h < H; ++h)
      if (lane + 64 * h < s) piv[ps[h]] = lane + 64 * h;
  }
  pr_for_square<SMAX>(s, [&](int r, int c) {  // rows at and beyond the rank
    if (r >= rank) F[r * ld + c] = 0.0;
  });
  __syncthreads();
  for (int u = tid; u < rank; u += NT) F[u * ld + piv[u]] = sqrt(dpv[u]);  // R(j, j)
  __syncthreads();
  // columns to position order: R(r, c) = F(r, piv[c]) for c >= r, zero below
  {
    constexpr int E = SMAX * SMAX / NT;
    double v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int idx = tid + k * NT, r = idx / SMAX, c = idx % SMAX;
      v[k] = (r < s && c < s && c >= r) ? F[r * ld + piv[c]] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int idx = tid + k * NT, r = idx / SMAX, c = idx % SMAX;
      if (r < s && c < s) F[r * ld + c] = v[k];
    }
    __syncthreads();
  }
  // filter (gpu/rbrp.cu:193-207): row energies, csum from the last row up
  for (int r = tid; r < s; r += NT) {
    double acc = 0.0;
#pragma unroll 8
    for (int c = r; c < s; ++c) acc += F[r * ld + c] * F[r * ld + c];
    sqr[r] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    // sums in serial order from the last row up, loads in batches of eight
    double total = 0.0;
    for (int u0 = 0; u0 < s; u0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = sqr[max(s - 1 - u0 - u, 0)];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u0 + u < s) total += v[u];
    }
    const double thr = total / (double)(filter_mode == 0 ? s : b);
    double csum = 0.0;
    int k = s;
    for (int u0 = 0; u0 < s && k == s; u0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = sqr[max(s - 1 - u0 - u, 0)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (k == s && u0 + u < s) {
          csum += v[u];
          if (csum >= thr) k = u0 + u;
        }
      }
    }
    int bt = s - k;
    if (rank == 0 || !(total > 0.0)) {
      bt = 0;
      bs->status = 1;
    }
    if (bt > rank) bt = rank;
    s_bt = bt;
  }
  __syncthreads();
  const int bt = s_bt;
  for (int u = tid; u < bt; u += NT) {
    const int jv = J[piv[u]];
    Jk[u] = jv;
    sj[u] = jv;
    if (Jkpos) Jkpos[u] = piv[u];
  }
  pr_for_square<SMAX>(bt, [&](int r, int c) {
    Rb[r + (int64_t)c * b] = (r <= c) ? F[r * ld + c] : 0.0;
  });
  __syncthreads();
  pr_upper_inverse<SMAX>(F, ld, bt, dinv, Rinv, b);
  // ascending sort of Jk by rank counting (entries are distinct); Jout flags
  for (int u = tid; u < bt; u += NT) {
    const int v = sj[u];
    int rk = 0;
    for (int q = 0; q < bt; ++q) rk += (sj[q] < v) ? 1 : 0;
    sorted[rk] = v;
    const int q0 = (int)(i + u);
    int member = 0;
    for (int q = 0; q < bt; ++q) member |= (sj[q] == q0) ? 1 : 0;
    flag[u] = !member;
  }
  __syncthreads();
  int nin = 0;
  for (int u = 0; u < bt; ++u) nin += ((int64_t)sorted[u] >= i + bt) ? 1 : 0;
  for (int u = tid; u < nin; u += NT) Jin[u] = sorted[bt - nin + u];
  for (int u = tid; u < bt; u += NT) {
    if (flag[u]) {
      int p = 0;
      for (int q = 0; q < u; ++q) p += flag[q];
      Jout[p] = (int)(i + u);
    }
  }
  if (tid == 0) {
    int nout = 0;
    for (int u = 0; u < bt; ++u) nout += flag[u];
    if (nin != nout) bs->status = 3;
    bs->bt = bt;
    bs->nswap = nin;
  }
}

// Second Cholesky-QR pass (the outputs of k_chol2_wave, for bt <= SMAX):
// G2 = R2^T R2 without pivoting, R2inv = R2^{-1}, Rout = R2 * Rb.
template <int SMAX>
__global__ __launch_bounds__(PrShape<SMAX>::NT) void k_chol2_regs(
    const double* __restrict__ G2, int64_t ldg, int b, RbrpBlockState* bs,
    const double* __restrict__ Rb, double* __restrict__ R2, double* __restrict__ R2inv,
    double* __restrict__ Rout, int64_t ldr) {
  constexpr int NT = PrShape<SMAX>::NT, TPR = PrShape<SMAX>::TPR;
  extern __shared__ double F[];  // bt x (bt + 1): G2, then R2 and its inverse
  __shared__ double lv[2][SMAX];
  __shared__ double dinv[SMAX], dpv[SMAX];
  __shared__ int s_bad;
  const int tid = threadIdx.x, a = tid / TPR, t = tid % TPR;
  const int bt = bs->bt;
  if (bt <= 0) return;
  const int ld = bt + 1;
  v16d T = pr_load_tiles<SMAX>(G2, ldg, bt, F, ld);
  if (tid == 0) s_bad = 0;
  __syncthreads();
  for (int j = 0; j < bt; ++j) {
    const int A = j >> 2, q = j & 3;
    // S(j, j) from the thread of tile (A, A): lane (A * TPR + A) % 64 of the
    // wave that holds row band A (the other waves read a value they ignore)
    const double d = pw_readlane(T[5 * q], (A * TPR + A) & 63);
    double* l = lv[j & 1];
    if (a == A) {
      const double dd = d > 0.0 ? d : 1.0;
      const double inv = pw_rsqrt(dd);
      if (t == A) {
        dpv[j] = dd;
        if (!(d > 0.0)) s_bad = 1;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = 4 * t + e;
        const double x = T[4 * q + e] * inv;
        const double v = (c <= j || c >= bt) ? 0.0 : x;
        l[c] = v;
        if (c < bt) F[j * ld + c] = v;  // (the diagonal entry is set after the loop)
      }
    }
    __syncthreads();
    T = pr_rank1<SMAX>(l, T);
  }
  for (int u = tid; u < bt; u += NT) F[u * ld + u] = sqrt(dpv[u]);
  __syncthreads();
  if (tid == 0 && s_bad) bs->status = 2;
  pr_for_square<SMAX>(bt, [&](int r, int c) {
    R2[r + (int64_t)c * b] = (r <= c) ? F[r * ld + c] : 0.0;
  });
  // Rout = R2 * Rb on the matrix core: block (I, J) = sum_{K=I}^{J} R2(I, K)
  // Rb(K, J).  Rb's upper triangle is staged in F's free strictly lower
  // triangle, transposed (Rb(r, c), r < c, at F[c * ld + r]), its diagonal in
  // dpv (R2's diagonal is final in F).  One wave per block.
  {
    __syncthreads();  // dpv (R2's pivots) consumed above
    {
      constexpr int E = SMAX * SMAX / NT;
      double v[E];  // every load in flight at once (clamped addresses)
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int idx = tid + k * NT, r = idx % SMAX, c = idx / SMAX;
        v[k] = Rb[min(r, bt - 1) + (int64_t)min(c, bt - 1) * b];
      }
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int idx = tid + k * NT, r = idx % SMAX, c = idx / SMAX;
        if (c < bt && r < c) F[c * ld + r] = v[k];
        if (c < bt && r == c) dpv[r] = v[k];
      }
    }
    __syncthreads();
    const int lane = tid & 63, j = lane & 15, g = lane >> 4, nw = NT >> 6;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nbk = (bt + 15) >> 4;
    for (int blk = w; blk < nbk * nbk; blk += nw) {  // wave-uniform
      const int I = blk / nbk, Jb = blk % nbk;
      v4d a4[2];
      a4[0] = a4[1] = (v4d){0.0, 0.0, 0.0, 0.0};
      for (int K = I; K <= Jb; ++K) {
        double av[4], bv[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int ka = 16 * K + 4 * s4 + g;
          const double u = pr_uload(F, ld, bt, 16 * I + j, ka);  // R2(I, K)
          av[s4] = (16 * I + j <= ka) ? u : 0.0;  // (F's lower triangle holds Rb)
          bv[s4] = pr_xload(F, ld, dpv, bt, ka, 16 * Jb + j);  // Rb(K, J), staged
        }
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
          a4[s4 & 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s4], bv[s4], a4[s4 & 1], 0, 0, 0);
      }
      const v4d acc = a4[0] + a4[1];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * I + 4 * v + g, c = 16 * Jb + j;
        if (r < bt && c < bt) Rout[r + (int64_t)c * ldr] = acc[v];
      }
    }
  }
  __syncthreads();  // the inverse overwrites F's lower triangle, read above
  pr_upper_inverse<SMAX>(F, ld, bt, dinv, R2inv, b);
}

template <int SMAX>
static hipError_t pr_allow(const void* fn, std::atomic<int>* cache) {
  return rbrp_allow_lds(fn, SMAX * (SMAX + 1) * (int)sizeof(double), cache);
}

hipError_t rbrp_launch_pivchol_regs(hipStream_t st, const double* G, int64_t ldg, int b,
                                    int filter_mode, int64_t i, const int* J, RbrpBlockState* bs,
                                    double* Rb, double* Rinv, int* Jk, int* Jin, int* Jout,
                                    int* Jkpos) {
  if (b < 1 || b > 128) return hipErrorInvalidValue;
  const size_t lds = (size_t)b * (size_t)(b + 1) * sizeof(double);
  static std::atomic<int> lds_set[RBRP_MAX_DEVICES];
  hipError_t e = pr_allow<128>((const void*)k_pivchol_regs<128>, lds_set);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pivchol_regs<128>, dim3(1), dim3(PrShape<128>::NT), lds, st, G, ldg, b,
                     filter_mode, i, J, bs, Rb, Rinv, Jk, Jin, Jout, Jkpos);
  return hipGetLastError();
}

hipError_t rbrp_launch_chol2_regs(hipStream_t st, const double* G2, int64_t ldg, int b,
                                  RbrpBlockState* bs, const double* Rb, double* R2, double* R2inv,
                                  double* Rout, int64_t ldr) {
  if (b < 1 || b > 128) return hipErrorInvalidValue;
  const size_t lds = (size_t)b * (size_t)(b + 1) * sizeof(double);
  static std::atomic<int> lds_set[RBRP_MAX_DEVICES];
  hipError_t e = pr_allow<128>((const void*)k_chol2_regs<128>, lds_set);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_chol2_regs<128>, dim3(1), dim3(PrShape<128>::NT), lds, st, G2, ldg, b, bs,
                     Rb, R2, R2inv, Rout, ldr);
  return hipGetLastError();
}

static size_t pw_lds_bytes() { return (size_t)PW_LDS_DOUBLES * sizeof(double); }

hipError_t rbrp_launch_upper_inv_blocks(hipStream_t st, int64_t k, const double* R, int64_t ldr,
                                        double* out) {
  if (k <= 0) return hipSuccess;
  const size_t lds = pw_lds_bytes();
  static std::atomic<int> lds_set[RBRP_MAX_DEVICES];
  hipError_t e = rbrp_allow_lds((const void*)k_upper_inv_blocks, (int)lds, lds_set);
  if (e != hipSuccess) return e;
  const unsigned nblk = (unsigned)(4 * ((k + 255) / 256));
  hipLaunchKernelGGL(k_upper_inv_blocks, dim3(nblk), dim3(64), lds, st, k, R, ldr, out);
  return hipGetLastError();
}

hipError_t rbrp_launch_pivchol_wave(hipStream_t st, const double* G, int64_t ldg, int b,
                                    int filter_mode, int64_t i, const int* J, RbrpBlockState* bs,
                                    double* Rb, double* Rinv, int* Jk, int* Jin, int* Jout,
                                    int* Jkpos) {
  if (b > PW_N) return hipErrorInvalidValue;
  const size_t lds = pw_lds_bytes();
  static std::atomic<int> lds_set[RBRP_MAX_DEVICES];
  hipError_t e = rbrp_allow_lds((const void*)k_pivchol_wave, (int)lds, lds_set);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pivchol_wave, dim3(1), dim3(64 * PW_NW), lds, st, G, ldg, b, filter_mode, i, J, bs,
                     Rb, Rinv, Jk, Jin, Jout, Jkpos);
  return hipGetLastError();
}

hipError_t rbrp_launch_chol2_wave(hipStream_t st, const double* G2, int64_t ldg, int b,
                                  RbrpBlockState* bs, const double* Rb, double* R2, double* R2inv,
                                  double* Rout, int64_t ldr) {
  if (b > PW_N) return hipErrorInvalidValue;
  const size_t lds = pw_lds_bytes();
  static std::atomic<int> lds_set[RBRP_MAX_DEVICES];
  hipError_t e = rbrp_allow_lds((const void*)k_chol2_wave, (int)lds, lds_set);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_chol2_wave, dim3(1), dim3(64 * PW_NW), lds, st, G2, ldg, b, bs, Rb, R2, R2inv,
                     Rout, ldr);
  return hipGetLastError();
}
