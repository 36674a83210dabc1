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
#include "rbrp_kernels.h"

#define KT 16

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
// Split reduction: dst(r, c) (=|+=) sum_{z=0}^{nsplit-1} P[z](r, c), fixed order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_splits(DimSpec dM, DimSpec dN,
                                                       const double* __restrict__ P, int64_t ldp,
                                                       int64_t pstride, int nsplit, MatRef dst,
                                                       int mode) {
  const int64_t M = dim_eval(dM), N = dim_eval(dN);
  const int64_t r = (int64_t)blockIdx.y * 256 + threadIdx.x;
  const int64_t c = blockIdx.x;
  if (r >= M || c >= N) return;
  double s = 0.0;
  for (int z = 0; z < nsplit; ++z) s += P[(int64_t)z * pstride + r + c * ldp];
  double* d = mat_col(dst, c) + r;
  if (mode == GEMM_ADD)
    *d += s;
  else if (mode == GEMM_SUB)
    *d -= s;
  else
    *d = s;
}

hipError_t rbrp_launch_reduce_splits(hipStream_t st, DimSpec M, DimSpec N, int64_t Mmax, int64_t Nmax,
                                     const double* P, int64_t ldp, int64_t pstride, int nsplit,
                                     MatRef dst, int mode) {
  if (Mmax <= 0 || Nmax <= 0) return hipSuccess;
  dim3 grid((unsigned)Nmax, (unsigned)((Mmax + 255) / 256));
  hipLaunchKernelGGL(k_reduce_splits, grid, dim3(256), 0, st, M, N, P, ldp, pstride, nsplit, dst,
                     mode);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// NN with update: C(r, c) = (mode) A(r, 0:K) B(0:K, c); optional per-tile
// squared column norms of the result.  The accumulator tile is staged through
// LDS so that C is read and written in 512-byte column runs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm_nn(int64_t Mr, DimSpec dN, DimSpec dK, MatRef A,
                                                 MatRef B, MatRef C, int mode,
                                                 double* __restrict__ normpart, int64_t ldnp) {
  const int64_t N = dim_eval(dN), K = dim_eval(dK);
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
  // stage the accumulators as Cs[c][r]
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wm * 32 + mi * 16 + mf_row(lane, v);
        const int c = wn * 32 + ni * 16 + mf_col(lane);
        Cs[c * 65 + r] = acc[mi][ni][v];
      }
  __syncthreads();
  // wave `wid` updates columns wid, wid+4, ...; lane = row within the tile
  const int64_t r = r0 + lane;
  const bool rok = r < Mr;
  for (int q = 0; q < 16; ++q) {
    const int c = wid + 4 * q;
    if (c0 + c >= N) break;  // uniform across the wave
    double* cp = mat_col(C, c0 + c);
    const double v = Cs[c * 65 + lane];
    double nv = 0.0;
    if (rok) {
      if (mode == GEMM_SUB)
        nv = cp[r] - v;
      else if (mode == GEMM_ADD)
        nv = cp[r] + v;
      else
        nv = v;
      cp[r] = nv;
    }
    if (normpart) {
      const double t = rbrp_wave_sum(nv * nv);
      if (lane == 0) normpart[(int64_t)blockIdx.y * ldnp + c0 + c] = t;
    }
  }
}

hipError_t rbrp_launch_gemm_nn(hipStream_t st, int64_t Mr, DimSpec N, int64_t Nmax, DimSpec K,
                               MatRef A, MatRef B, MatRef C, int mode, double* normpart,
                               int64_t ldnp) {
  if (Mr <= 0 || Nmax <= 0) return hipSuccess;
  dim3 grid((unsigned)((Nmax + 63) / 64), (unsigned)((Mr + 63) / 64));
  hipLaunchKernelGGL(k_gemm_nn, grid, dim3(256), 0, st, Mr, N, K, A, B, C, mode, normpart, ldnp);
  return hipGetLastError();
}
