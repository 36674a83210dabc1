// RBRP loop kernels other than the GEMMs: the HBM-bound norm pass, the
// weighted sampler, the LDS-resident pivoted Cholesky + robust filter of the
// sampled block, CholQR2 re-factorisation, the symmetric-difference
// permutation, residual sums and the diagonal-block solve of X = R11^{-1} R12.
#include <algorithm>
#include <climits>

#include "rbrp_kernels.h"

// ---------------------------------------------------------------------------
// Norm pass: Aw(:, j) = [A(0:m, j); 0 (pad rows m..mp)], diags[j] = ||A(:,j)||^2.
// One workgroup per column; loads are coalesced along the column.
// Reference: column_norm_squared (gpu/rbrp.cu:82-99, called at :354) plus the
// working copy at gpu/rbrp.cu:337-338.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy_colnorm(const double* __restrict__ A, int64_t lda,
                                                      int64_t m, int64_t mp,
                                                      double* __restrict__ Aw,
                                                      double* __restrict__ diags) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  const double* a = A + j * lda;
  double* w = Aw + j * mp;
  double acc0 = 0.0, acc1 = 0.0;
  int64_t r = threadIdx.x;
  for (; r + 256 < m; r += 512) {
    const double v0 = a[r], v1 = a[r + 256];
    w[r] = v0;
    w[r + 256] = v1;
    acc0 = fma(v0, v0, acc0);
    acc1 = fma(v1, v1, acc1);
  }
  for (; r < mp; r += 256) {
    const double v = r < m ? a[r] : 0.0;
    w[r] = v;
    acc0 = fma(v, v, acc0);
  }
  const double t = rbrp_block_sum(acc0 + acc1, red);
  if (threadIdx.x == 0) diags[j] = t;
}

hipError_t rbrp_launch_copy_colnorm(hipStream_t st, const double* A, int64_t lda, int64_t m,
                                    int64_t mp, int64_t n, double* Aw, double* diags) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_colnorm, dim3((unsigned)n), dim3(256), 0, st, A, lda, m, mp, Aw, diags);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Norm pass without a working copy: diags[j] = ||A(:, j)||^2, a pure read of
// A (the FP64 loop's first block reads A itself; see k_gemm_tn64 / the update
// kernels' first-block forms).  One workgroup per column, each lane streaming
// 16-byte loads (four in flight) when the column starts on a 16-byte boundary
// (A aligned and lda even), else single doubles.  The per-lane sums are added
// in a fixed order, then the workgroup sum.
// Reference: column_norm_squared (gpu/rbrp.cu:82-99, called at :354).
// ---------------------------------------------------------------------------
template <bool VEC>
__global__ __launch_bounds__(256) void k_colnorm(const double* __restrict__ A, int64_t lda,
                                                 int64_t m, double* __restrict__ diags) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  const double* a = A + j * lda;
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
  int64_t r;
  if (VEC) {
    // nontemporal loads: A is streamed once here (measured 7.0 vs 6.1 TB/s
    // for the same loop with ordinary loads on a 1.6 GB matrix)
    const v2d* a2 = (const v2d*)a;
    const int64_t m2 = m >> 1;  // whole pairs
    int64_t q = threadIdx.x;
    for (; q + 768 < m2; q += 1024) {
      const v2d v0 = __builtin_nontemporal_load(a2 + q), v1 = __builtin_nontemporal_load(a2 + q + 256),
                v2 = __builtin_nontemporal_load(a2 + q + 512), v3 = __builtin_nontemporal_load(a2 + q + 768);
      acc0 = fma(v0.x, v0.x, acc0);
      acc1 = fma(v0.y, v0.y, acc1);
      acc2 = fma(v1.x, v1.x, acc2);
      acc3 = fma(v1.y, v1.y, acc3);
      acc0 = fma(v2.x, v2.x, acc0);
      acc1 = fma(v2.y, v2.y, acc1);
      acc2 = fma(v3.x, v3.x, acc2);
      acc3 = fma(v3.y, v3.y, acc3);
    }
    for (; q < m2; q += 256) {
      const v2d v0 = __builtin_nontemporal_load(a2 + q);
      acc0 = fma(v0.x, v0.x, acc0);
      acc1 = fma(v0.y, v0.y, acc1);
    }
    r = 2 * m2 + threadIdx.x;  // the odd last row
  } else {
    r = threadIdx.x;
    for (; r + 768 < m; r += 1024) {
      const double v0 = a[r], v1 = a[r + 256], v2 = a[r + 512], v3 = a[r + 768];
      acc0 = fma(v0, v0, acc0);
      acc1 = fma(v1, v1, acc1);
      acc2 = fma(v2, v2, acc2);
      acc3 = fma(v3, v3, acc3);
    }
  }
  for (; r < m; r += 256) acc0 = fma(a[r], a[r], acc0);
  const double t = rbrp_block_sum((acc0 + acc1) + (acc2 + acc3), red);
  if (threadIdx.x == 0) diags[j] = t;
}

hipError_t rbrp_launch_colnorm(hipStream_t st, const double* A, int64_t lda, int64_t m, int64_t n,
                               double* diags) {
  if (n <= 0) return hipSuccess;
  const bool vec = ((uintptr_t)A % 16 == 0) && (lda % 2 == 0);
  if (vec)
    hipLaunchKernelGGL(k_colnorm<true>, dim3((unsigned)n), dim3(256), 0, st, A, lda, m, diags);
  else
    hipLaunchKernelGGL(k_colnorm<false>, dim3((unsigned)n), dim3(256), 0, st, A, lda, m, diags);
  return hipGetLastError();
}

// V(:, p) = [A(0:m, J[p]); 0 (rows m..mp)] for p < bs->s: the sampled columns
// of the FP64 loop's first block, read from A (no working copy exists yet).
__global__ __launch_bounds__(256) void k_gather_first(const double* __restrict__ A, int64_t lda,
                                                      int64_t m, int64_t mp,
                                                      const RbrpBlockState* __restrict__ bs,
                                                      const int* __restrict__ J,
                                                      double* __restrict__ V, int64_t ldv) {
  const int p = blockIdx.y;
  if (p >= bs->s) return;
  const double* a = A + (int64_t)J[p] * lda;
  double* v = V + (int64_t)p * ldv;
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < mp; r += (int64_t)gridDim.x * 256)
    v[r] = r < m ? a[r] : 0.0;
}

hipError_t rbrp_launch_gather_first(hipStream_t st, const double* A, int64_t lda, int64_t m,
                                    int64_t mp, int b, const RbrpBlockState* bs, const int* J,
                                    double* V, int64_t ldv) {
  if (b <= 0 || mp <= 0) return hipSuccess;
  const int64_t bx = std::min<int64_t>((mp + 255) / 256, 64);
  hipLaunchKernelGGL(k_gather_first, dim3((unsigned)bx, (unsigned)b), dim3(256), 0, st, A, lda, m, mp,
                     bs, J, V, ldv);
  return hipGetLastError();
}

// Sum of x[0:n) into the block state (which = 0: N0).
__global__ __launch_bounds__(1024) void k_sum_to_state(const double* __restrict__ x, int64_t n,
                                                       RbrpBlockState* bs, int which) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 1024) acc += x[j];
  const double t = rbrp_block_sum(acc, red);
  if (threadIdx.x == 0) {
    if (which == 0) {
      bs->N0 = t;
      bs->draws = 0ull;
      bs->status = 0;
      bs->s = 0;
      bs->bt = 0;
      bs->nswap = 0;
      bs->resid = t;
      bs->err = (t > 0.0) ? 1.0 : 0.0;
    }
  }
}

hipError_t rbrp_launch_sum_to_state(hipStream_t st, const double* x, int64_t n, RbrpBlockState* bs,
                                    int which) {
  hipLaunchKernelGGL(k_sum_to_state, dim3(1), dim3(1024), 0, st, x, n, bs, which);
  return hipGetLastError();
}

// p[0:count) = 0 (a kernel rather than a memset node, so a captured graph
// replays it like any other launch).
__global__ __launch_bounds__(256) void k_zero(double* __restrict__ p, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < count; j += stride) p[j] = 0.0;
}

hipError_t rbrp_launch_zero(hipStream_t st, double* p, int64_t count) {
  if (count <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((count + 255) / 256, 2048);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, count);
  return hipGetLastError();
}

// *out = sum of x[0:n) (one workgroup, fixed order).
__global__ __launch_bounds__(1024) void k_sum(const double* __restrict__ x, int64_t n, double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 1024) acc += x[j];
  const double t = rbrp_block_sum(acc, red);
  if (threadIdx.x == 0) *out = t;
}

hipError_t rbrp_launch_sum(hipStream_t st, const double* x, int64_t n, double* out) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, st, x, n, out);
  return hipGetLastError();
}

__global__ void k_iota(int* perm, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < n) perm[j] = (int)j;
}

hipError_t rbrp_launch_iota(hipStream_t st, int* perm, int64_t n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, perm, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Weighted sampling without replacement (one workgroup).
//
// s = min(b, kmax - i, nnz(diags[i:n])) distinct positions of [i, n) with
// weights diags[i:n], by the rounds algorithm of NumPy's
// Generator.choice(replace=False, p) (the Python reference's sampler,
// RBRP.ipynb c11:108): each round draws s - found uniforms, zeroes the weights
// found so far, forms cdf = inclusive_scan(w) / cdf[N-1], maps each uniform to
// the first index with cdf > u, and appends the new distinct picks in order of
// first occurrence.  Replaces select_indices (gpu/rbrp.cu:102-146: Thrust
// scan + upper_bound + sort + unique, with replacement).
// ---------------------------------------------------------------------------
__device__ inline double block_exclusive_scan(double v, double* wsum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  double ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = 0.0;
  __syncthreads();
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  // off = wsum[0] + ... + wsum[wid - 1] in order (the +0.0 of the later
  // waves leaves it unchanged; the 16 reads issue together)
  double off = 0.0;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const double y = wsum[q < nw ? q : 0];
    off += (q < wid) ? y : 0.0;
  }
  return off + ex;
}

#define SAMPLE_MAXB 256

// First occurrences of a round's picks.  Appends to found[nf0:] the values of
// newidx[0:r] that are not in found[0:nf0], each once, in order of first
// occurrence, and returns the new count to every thread (nf0 + r <=
// SAMPLE_MAXB; 1024-thread workgroup).  An LDS open-addressing table maps
// every value to the smallest draw holding it (-1 for the found ones); draw d
// is new when it owns its value.  The positions follow from a ballot per wave
// and the wave counts.
#define COLLECT_SLOTS 512
struct CollectLds {
  int key[COLLECT_SLOTS];
  int own[COLLECT_SLOTS];
  int wcnt[16];
};

__device__ inline int collect_slot(CollectLds& c, int v, int d) {
  unsigned h = ((unsigned)v * 2654435761u) >> 23;  // 9 bits
  // at most SAMPLE_MAXB distinct values in COLLECT_SLOTS slots: a free or
  // matching slot is always found
  while (true) {
    const int prev = atomicCAS(&c.key[h], -1, v);
    if (prev == -1 || prev == v) {
      atomicMin(&c.own[h], d);
      return (int)h;
    }
    h = (h + 1) & (COLLECT_SLOTS - 1);
  }
}

__device__ inline int collect_first(const int* newidx, int r, int* found, int nf0, CollectLds& c) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int q = tid; q < COLLECT_SLOTS; q += 1024) {
    c.key[q] = -1;
    c.own[q] = INT_MAX;
  }
  __syncthreads();
  if (tid < nf0) (void)collect_slot(c, found[tid], -1);
  int slot = 0;
  if (tid < r) slot = collect_slot(c, newidx[tid], tid);
  __syncthreads();
  const bool isnew = tid < r && c.own[slot] == tid;
  const unsigned long long m = __ballot(isnew);
  const int below = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) c.wcnt[wid] = __popcll(m);
  __syncthreads();
  int off = nf0, tot = nf0;
  for (int q = 0; q < 16; ++q) {
    off += (q < wid) ? c.wcnt[q] : 0;
    tot += c.wcnt[q];
  }
  if (isnew) found[off + below] = newidx[tid];
  __syncthreads();
  return tot;
}

// Weights and CDF live in LDS when the candidate range fits (N <= SAMPLE_LDS_N),
// otherwise in the global scratch arrays w / cdf.
#define SAMPLE_LDS_N 8192

// The block's first column is i0 + *i_dyn (i_dyn optional).  With check_prev
// the state still holds the previous block's results, and the call is a
// no-op (s = 0, so the following kernels of the block do nothing) when that
// block ended the loop: a numerical failure, an empty block, the squared
// error at or below tol2, or the rank cap reached.  in_lds says whether the
// launch reserved the LDS arrays.
__global__ __launch_bounds__(1024) void k_sample(const double* __restrict__ diags, int64_t i0,
                                                 const int* i_dyn, int check_prev, double tol2,
                                                 int in_lds_ok, int64_t n, int64_t kmax, int b,
                                                 uint64_t key0, uint64_t key1, uint64_t c0,
                                                 uint64_t c1, uint64_t c2, uint64_t c3,
                                                 double* __restrict__ gw,
                                                 double* __restrict__ gcdf, RbrpBlockState* bs,
                                                 int* __restrict__ J) {
  extern __shared__ double lds[];
  __shared__ int found[SAMPLE_MAXB];
  __shared__ int newidx[SAMPLE_MAXB];
  __shared__ CollectLds cl;
  __shared__ int ired[16];
  __shared__ double wsum[16];
  __shared__ int s_done;
  const int tid = threadIdx.x;
  const int64_t i = i0 + (i_dyn ? (int64_t)*i_dyn : 0);
  if (check_prev) {
    if (tid == 0)
      s_done = bs->status != 0 || bs->s <= 0 || bs->bt <= 0 || bs->err <= tol2 || i >= kmax;
    __syncthreads();
    if (s_done) {
      if (tid == 0) bs->s = 0;
      return;
    }
  }
  const int64_t N = n - i;
  const bool in_lds = in_lds_ok && N <= SAMPLE_LDS_N;
  double* w = in_lds ? lds : gw;
  double* cdf = in_lds ? lds + SAMPLE_LDS_N : gcdf;
  int cnt = 0;
  for (int64_t j = tid; j < N; j += 1024) {
    double v = diags[i + j];
    v = (v > 0.0) ? v : 0.0;
    w[j] = v;
    cnt += (v > 0.0) ? 1 : 0;
  }
  const int nnz = rbrp_block_sum_int(cnt, ired);
  int64_t s64 = (int64_t)b;
  if (kmax - i < s64) s64 = kmax - i;
  if ((int64_t)nnz < s64) s64 = nnz;
  const int s = (int)(s64 > 0 ? s64 : 0);
  unsigned long long draws = bs->draws;
  if (s <= 0) {
    if (tid == 0) {
      bs->s = 0;
      bs->nnz = nnz;
    }
    return;
  }
  int nfound = 0;
  const uint64_t base[4] = {c0, c1, c2, c3};
  const int64_t chunk = (N + 1023) / 1024;
  const int64_t j0 = min(N, (int64_t)tid * chunk);
  const int64_t j1 = min(N, j0 + chunk);
  // every round adds at least one new index (the first draw of a round can
  // only land on a positive, not yet found weight), so s rounds suffice
  for (int round = 0; round <= s; ++round) {
    const int nf = nfound;
    if (nf >= s) break;
    const int r = s - nf;
    for (int t = tid; t < nf; t += 1024) w[found[t]] = 0.0;
    __syncthreads();
    double loc = 0.0;
    for (int64_t j = j0; j < j1; ++j) loc += w[j];
    double run = block_exclusive_scan(loc, wsum);
    for (int64_t j = j0; j < j1; ++j) {
      run += w[j];
      cdf[j] = run;
    }
    __syncthreads();
    const double total = cdf[N - 1];
    if (tid < r) {
      const double u = rbrp_u64_to_unit(rbrp_philox_word(base, key0, key1, draws + tid));
      int64_t lo = 0, hi = N;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] / total > u)
          hi = mid;
        else
          lo = mid + 1;
      }
      if (lo >= N) lo = N - 1;
      newidx[tid] = (int)lo;
    }
    __syncthreads();
    // keep the first occurrence of each index not found before
    nfound = collect_first(newidx, r, found, nf, cl);
    draws += (unsigned long long)r;
  }
  if (nfound < s) {
    if (tid == 0) {
      bs->s = 0;
      bs->status = 4;
    }
    return;
  }
  for (int t = tid; t < s; t += 1024) J[t] = (int)(i + found[t]);
  if (tid == 0) {
    bs->s = s;
    bs->draws = draws;
    bs->nnz = nnz;
  }
}

// ---------------------------------------------------------------------------
// Wide sampler: the same rounds algorithm for a candidate range N larger than
// SAMPLE_WIDE_N, over every compute unit.  The weights are cut into tiles of
// WIDE_TILE (thread k of a 1024-thread workgroup owns the tile's elements 4k
// .. 4k+3).  The CDF value of element j of tile t is defined as
//   cdf[j] = P[t] + run(j),
// run(j) being the inclusive running sum inside the tile in one fixed order
// (thread k starts from the exclusive prefix of the 1024 thread sums and adds
// its four weights in turn) and P[t] the tile prefixes, with the total =
// P[last] + S[last] = cdf[N - 1] (S[t] = run of the tile's last element).
// Every kernel that needs run() computes it with the same device function, so
// the tile sums, the tile search and the in-tile search agree exactly and the
// CDF is non-decreasing.  The CDF is never stored: one pass reads the weights
// (8 bytes per candidate) and writes S and the non-zero counts per tile; one
// workgroup forms P and the total, draws the round's uniforms and picks each
// draw's tile; one workgroup per draw re-scans its tile to find the index;
// one workgroup keeps the first occurrences and runs any further round itself
// (the tiles holding found indices re-summed with those weights zeroed).
// ---------------------------------------------------------------------------
#define WIDE_TILE 4096
#define SAMPLE_WIDE_N 65536

struct WideHdr {
  int64_t i, N, ntiles;
  int s, r, nnz, done;
  double total;
  unsigned long long draws;
};

__device__ inline bool wide_found(const int* fs, int nf, int64_t j) {
  int lo = 0, hi = nf;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)fs[mid] < j)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < nf && (int64_t)fs[lo] == j;
}

// This thread's four weights of tile t (elements 4k .. 4k+3 for thread k),
// clipped at zero, with the fs[0:nf] sorted found positions zeroed (N >= 1).
__device__ inline void wide_load(const double* __restrict__ diags, int64_t i, int64_t N, int64_t t,
                                 const int* fs, int nf, double w[4]) {
  // (the loads are unconditional, at positions clamped into [0, N), so that
  // they issue back to back; out-of-range positions get weight zero after)
  const int64_t j0 = t * WIDE_TILE + 4 * (int64_t)threadIdx.x;
  double v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = diags[i + min(j0 + q, N - 1)];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t j = j0 + q;
    double x = (j < N && v[q] > 0.0) ? v[q] : 0.0;
    if (nf > 0 && x > 0.0 && wide_found(fs, nf, j)) x = 0.0;
    w[q] = x;
  }
}

// Inclusive running sums of this thread's four weights (from wide_load) in the
// tile's fixed order, and its count of positive weights.  wsum: 16 doubles.
__device__ inline void wide_scan(const double w[4], double* wsum, double run[4], int* cnt) {
  double loc = 0.0;
  int c = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    loc += w[q];
    c += (w[q] > 0.0) ? 1 : 0;
  }
  double r = block_exclusive_scan(loc, wsum);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    r += w[q];
    run[q] = r;
  }
  *cnt = c;
}

// wide_load + wide_scan of tile t.
__device__ inline void wide_tile(const double* __restrict__ diags, int64_t i, int64_t N, int64_t t,
                                 const int* fs, int nf, double* wsum, double run[4], int* cnt) {
  double w[4];
  wide_load(diags, i, N, t, fs, nf, w);
  wide_scan(w, wsum, run, cnt);
}

// Scratch of the wide sampler, carved from the sampler's w / cdf arrays.
struct WideScratch {
  double *S, *P, *u, *tp;  // tp[d]: P of draw d's tile
  int *tnnz, *tilesel, *newidx;
  WideHdr* hdr;
};

__host__ __device__ inline WideScratch wide_scratch(double* gw, double* gcdf, int64_t ntiles_max) {
  WideScratch w;
  w.S = gw;
  w.P = gw + ntiles_max;
  w.hdr = (WideHdr*)gcdf;
  w.u = gcdf + 8;
  w.tp = gcdf + 8 + SAMPLE_MAXB;
  w.tilesel = (int*)(gcdf + 8 + 2 * SAMPLE_MAXB);
  w.newidx = w.tilesel + SAMPLE_MAXB;
  w.tnnz = w.newidx + SAMPLE_MAXB;
  return w;
}

__device__ inline int64_t wide_first_col(int64_t i0, const int* i_dyn) {
  return i0 + (i_dyn ? (int64_t)*i_dyn : 0);
}

// Tile pass: S[t] and the positive-weight count of every tile, the
// workgroups striding over the tiles.  One barrier pair per tile: the wave
// totals and the wave counts (ballots) go to LDS together.  S[t] follows
// wide_scan's arithmetic for the tile's last element (the sum of waves 0 ..
// 14, then the last wave's exclusive prefix, then the thread's four weights
// in turn).  wide_load's loads are plain ones: the nontemporal form of these
// 8-byte loads ran at two thirds of the rate.
__global__ __launch_bounds__(1024) void k_wide_tiles(const double* __restrict__ diags, int64_t i0,
                                                     const int* i_dyn, int64_t n, WideScratch ws) {
  __shared__ double wsum[16];
  __shared__ int ired[16];
  const int64_t i = wide_first_col(i0, i_dyn);
  const int64_t N = n - i;
  const int64_t ntiles = N > 0 ? (N + WIDE_TILE - 1) / WIDE_TILE : 0;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // full tiles of a 16-byte aligned range load as two 16-byte units a thread
  const bool vec = (((uintptr_t)(diags + i)) & 15) == 0;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    double w[4];
    if (vec && t + 1 < ntiles) {
      typedef double v2d __attribute__((ext_vector_type(2)));
      const v2d* p = (const v2d*)(diags + i + t * WIDE_TILE + 4 * (int64_t)tid);
      const v2d a = p[0], b = p[1];
      w[0] = a.x > 0.0 ? a.x : 0.0;
      w[1] = a.y > 0.0 ? a.y : 0.0;
      w[2] = b.x > 0.0 ? b.x : 0.0;
      w[3] = b.y > 0.0 ? b.y : 0.0;
    } else {
      wide_load(diags, i, N, t, nullptr, 0, w);
    }
    double x = 0.0;
    int c = 0;  // the wave's count of positive weights
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      x += w[q];
      c += __popcll(__ballot(w[q] > 0.0));
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    double ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = 0.0;
    __syncthreads();
    if (lane == 63) wsum[wid] = x;
    if (lane == 0) ired[wid] = c;
    __syncthreads();
    if (tid == 1023) {
      double r = 0.0;
      for (int q = 0; q < 15; ++q) r += wsum[q];
      r += ex;
#pragma unroll
      for (int q = 0; q < 4; ++q) r += w[q];
      ws.S[t] = r;
    } else if (tid == 0) {
      int cnt = 0;
      for (int q = 0; q < 16; ++q) cnt += ired[q];
      ws.tnnz[t] = cnt;
    }
  }
}

// LDS of the pick and finish kernels when the launch reserves it (E !=
// nullptr): the tile end values P[t] + S[t], thread k's run of per tiles at
// E[k (per + 1) + q], with the run's starting prefix in the run's extra slot
// E[k (per + 1) + per] (the odd stride also keeps the per-thread walks free
// of bank conflicts).
struct WideLds {
  double* E;
  int per;       // tiles per run, ceil(ntiles / 1024)
  uint64_t inv;  // ceil(2^32 / per): t / per = (t inv) >> 32 for t < 2^15
};

__host__ __device__ inline int64_t wide_lds_doubles(int64_t ntiles) {
  return 1024 * ((ntiles + 1023) / 1024 + 1);
}

// The LDS path holds per <= WIDE_MAX_PER (ntiles <= WIDE_LDS_MAX_TILES <
// 2^15), where (t ceil(2^32 / per)) >> 32 = floor(t / per) exactly (the
// rounding error t (inv - 2^32 / per) / 2^32 < 2^-17 stays below 1 / per).
#define WIDE_MAX_PER 20
#define WIDE_LDS_MAX_TILES (WIDE_MAX_PER * 1024)

// LDS slot of tile t.
__device__ inline int wide_slot(const WideLds& L, int t) {
  const int k = (int)(((uint64_t)t * L.inv) >> 32);
  return k * (L.per + 1) + (t - k * L.per);
}

// Loads S[t] (and, with cnt, adds tnnz[t]) for this thread's tiles t = tid +
// 1024 r into their LDS slots: coalesced loads, 8 tiles in flight.
__device__ inline void wide_stage(const WideScratch& ws, int64_t ntiles, const WideLds& L, int* cnt) {
  const int tid = threadIdx.x;
  for (int t = tid; t < (int)ntiles; t += 8 * 1024) {
    // (loads at clamped positions, so that all issue back to back)
    double v[8];
    int c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int tq = min(t + 1024 * q, (int)ntiles - 1);
      v[q] = ws.S[tq];
      c[q] = cnt ? ws.tnnz[tq] : 0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (t + 1024 * q < (int)ntiles) {
        L.E[wide_slot(L, t + 1024 * q)] = v[q];
        if (cnt) *cnt += c[q];
      }
    }
  }
}

// P[t] (exclusive tile prefixes) and the total from S.  Thread k owns the run
// of tiles [k per, (k + 1) per) (per = ceil(ntiles / 1024)): it sums the run
// in order, takes the workgroup's exclusive scan of the run sums, and walks
// the run again adding S[t] to that prefix (P[t] is the value before the
// addition, P[t] + S[t] the value after).  With L.E the walks run in LDS (S
// staged there by wide_stage, here unless `staged`), and the end values and
// the run prefixes stay there for wide_pick_tile and wide_P; without it they
// read S from global memory and write P there.  The arithmetic is the same
// either way.  Every thread returns the total.
__device__ inline double wide_prefix(const WideScratch& ws, int64_t ntiles, const WideLds& L,
                                     bool staged, double* wsum, double* bcast) {
  const int tid = threadIdx.x;
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t t0 = min(ntiles, (int64_t)tid * per), t1 = min(ntiles, t0 + per);
  if (L.E) {
    const int ld = L.per + 1;
    if (!staged) {
      wide_stage(ws, ntiles, L, nullptr);
      __syncthreads();
    }
    // (the run read from LDS at once; P stays in LDS, read by wide_P)
    double* e = L.E + tid * ld;
    const int cnt = (int)(t1 - t0);
    // (slot min(q, per) stays inside the thread's run; the values past cnt
    // become +0.0, which leaves the sums unchanged)
    double v[WIDE_MAX_PER];
#pragma unroll
    for (int q = 0; q < WIDE_MAX_PER; ++q) v[q] = e[q < L.per ? q : L.per];
#pragma unroll
    for (int q = 0; q < WIDE_MAX_PER; ++q) v[q] = (q < cnt) ? v[q] : 0.0;
    double loc = 0.0;
#pragma unroll
    for (int q = 0; q < WIDE_MAX_PER; ++q) loc += v[q];
    double run = block_exclusive_scan(loc, wsum);
    e[L.per] = run;
#pragma unroll
    for (int q = 0; q < WIDE_MAX_PER; ++q) {
      if (q < cnt) {
        run += v[q];
        e[q] = run;
      }
    }
    if (cnt > 0 && t1 == ntiles) *bcast = run;  // P[last] + S[last]
    (void)ws;
  } else {
    // (eight loads issued together per step; adding the +0.0 of a missing
    // tile leaves every sum unchanged, so the order of the additions is the
    // plain left-to-right one)
    double loc = 0.0;
    for (int64_t t = t0; t < t1; t += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (t + q < t1) ? ws.S[t + q] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) loc += v[q];
    }
    double run = block_exclusive_scan(loc, wsum);
    for (int64_t t = t0; t < t1; t += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (t + q < t1) ? ws.S[t + q] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (t + q < t1) ws.P[t + q] = run;
        run += v[q];
      }
    }
    if (t0 < t1 && t1 == ntiles) *bcast = run;  // P[last] + S[last]
  }
  // (P is read back within this workgroup only, or by later kernels: the
  // barrier's workgroup-scope ordering suffices)
  __syncthreads();
  return *bcast;
}

// P[t] after wide_prefix (from LDS when it left the values there).
__device__ inline double wide_P(const WideScratch& ws, const WideLds& L, int64_t t) {
  if (!L.E) return ws.P[t];
  const int k = (int)(((uint64_t)t * L.inv) >> 32), q = (int)t - k * L.per;
  return L.E[k * (L.per + 1) + (q ? q - 1 : L.per)];
}

// First tile t with (P[t] + S[t]) / total > u (the end values from LDS when
// wide_prefix left them there).
__device__ inline int64_t wide_pick_tile(const WideScratch& ws, const WideLds& L, int64_t ntiles,
                                         double total, double u) {
  int64_t lo = 0, hi = ntiles;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const double e = L.E ? L.E[wide_slot(L, (int)mid)] : ws.P[mid] + ws.S[mid];
    if (e / total > u)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo < ntiles ? lo : ntiles - 1;
}

// First position j of tile t with (P[t] + run(j)) / total > u (all threads
// call; every thread returns it).
__device__ inline int64_t wide_find(const double* __restrict__ diags, int64_t i, int64_t N,
                                    double Pt, int64_t t, double total, double u, const int* fs,
                                    int nf, double* wsum, int* imin) {
  double run[4];
  int c;
  wide_tile(diags, i, N, t, fs, nf, wsum, run, &c);
  int best = INT_MAX;
#pragma unroll
  for (int q = 3; q >= 0; --q) {
    const int64_t j = t * WIDE_TILE + 4 * (int64_t)threadIdx.x + q;
    if (j < N && (Pt + run[q]) / total > u) best = (int)(j - t * WIDE_TILE);
  }
  // workgroup minimum
  for (int o = 32; o > 0; o >>= 1) best = min(best, __shfl_xor(best, o, RBRP_WAVE));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) imin[threadIdx.x >> 6] = best;
  __syncthreads();
  int m = INT_MAX;
  for (int q = 0; q < 16; ++q) m = min(m, imin[q]);
  if (m == INT_MAX) return min(N, (t + 1) * WIDE_TILE) - 1;
  return t * WIDE_TILE + m;
}

__device__ inline WideLds wide_lds(double* lds, int lds_ok, int64_t ntiles) {
  WideLds L;
  L.E = (lds_ok && ntiles <= WIDE_LDS_MAX_TILES) ? lds : nullptr;  // (per <= WIDE_MAX_PER)
  L.per = (int)((ntiles + 1023) / 1024);
  L.inv = L.E ? (0x100000000ull + (uint64_t)L.per - 1) / (uint64_t)L.per : 0ull;
  return L;
}

// Round 1: s, the tile prefixes, the total, the round's uniforms and their
// tiles.  lds_ok: the launch reserved wide_lds_doubles(ntiles) doubles of LDS.
__global__ __launch_bounds__(1024) void k_wide_pick(int64_t i0, const int* i_dyn, int check_prev,
                                                    double tol2, int64_t n, int64_t kmax, int b,
                                                    uint64_t key0, uint64_t key1, uint64_t c0,
                                                    uint64_t c1, uint64_t c2, uint64_t c3,
                                                    int lds_ok, WideScratch ws, RbrpBlockState* bs) {
  extern __shared__ double wide_dyn[];
  __shared__ double wsum[16];
  __shared__ int ired[16];
  __shared__ double bcast;
  __shared__ int s_done;
  const int tid = threadIdx.x;
  const int64_t i = wide_first_col(i0, i_dyn);
  if (tid == 0) {
    s_done = check_prev && (bs->status != 0 || bs->s <= 0 || bs->bt <= 0 || bs->err <= tol2 || i >= kmax);
    if (s_done) {
      bs->s = 0;
      ws.hdr->done = 1;
    }
  }
  __syncthreads();
  if (s_done) return;
  const int64_t N = n - i;
  const int64_t ntiles = N > 0 ? (N + WIDE_TILE - 1) / WIDE_TILE : 0;
  const WideLds L = wide_lds(wide_dyn, lds_ok, ntiles);
  int c = 0;
  if (L.E) {
    wide_stage(ws, ntiles, L, &c);  // (the barrier in the count's sum orders the LDS writes)
  } else {
    for (int64_t t = tid; t < ntiles; t += 8 * 1024) {
      int v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (t + 1024 * q < ntiles) ? ws.tnnz[t + 1024 * q] : 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) c += v[q];
    }
  }
  const int nnz = rbrp_block_sum_int(c, ired);
  int64_t s64 = (int64_t)b;
  if (kmax - i < s64) s64 = kmax - i;
  if ((int64_t)nnz < s64) s64 = nnz;
  const int s = (int)(s64 > 0 ? s64 : 0);
  if (s <= 0) {
    if (tid == 0) {
      bs->s = 0;
      bs->nnz = nnz;
      ws.hdr->done = 1;
    }
    return;
  }
  const double total = wide_prefix(ws, ntiles, L, true, wsum, &bcast);
  const unsigned long long draws = bs->draws;
  const uint64_t base[4] = {c0, c1, c2, c3};
  if (tid < s) {
    const double u = rbrp_u64_to_unit(rbrp_philox_word(base, key0, key1, draws + tid));
    const int64_t t = wide_pick_tile(ws, L, ntiles, total, u);
    ws.u[tid] = u;
    ws.tilesel[tid] = (int)t;
    ws.tp[tid] = wide_P(ws, L, t);
  }
  if (tid == 0) {
    WideHdr* h = ws.hdr;
    h->i = i;
    h->N = N;
    h->ntiles = ntiles;
    h->s = s;
    h->r = s;
    h->nnz = nnz;
    h->done = 0;
    h->total = total;
    h->draws = draws;
  }
}

// Round 1's in-tile searches: workgroup d resolves draw d.
__global__ __launch_bounds__(1024) void k_wide_resolve(const double* __restrict__ diags,
                                                       WideScratch ws) {
  __shared__ double wsum[16];
  __shared__ int imin[16];
  const WideHdr* h = ws.hdr;
  const int d = blockIdx.x;
  if (h->done || d >= h->r) return;  // uniform
  const int64_t j = wide_find(diags, h->i, h->N, ws.tp[d], ws.tilesel[d], h->total, ws.u[d], nullptr,
                              0, wsum, imin);
  if (threadIdx.x == 0) ws.newidx[d] = (int)j;
}

// Keeps round 1's first occurrences, runs the later rounds, writes J and the
// state (lds_ok as for k_wide_pick).
__global__ __launch_bounds__(1024) void k_wide_finish(const double* __restrict__ diags, int64_t kmax,
                                                      uint64_t key0, uint64_t key1, uint64_t c0,
                                                      uint64_t c1, uint64_t c2, uint64_t c3,
                                                      int lds_ok, WideScratch ws, RbrpBlockState* bs,
                                                      int* __restrict__ J) {
  extern __shared__ double wide_dyn[];
  __shared__ double wsum[16];
  __shared__ int imin[16];
  __shared__ double bcast;
  __shared__ int found[SAMPLE_MAXB], fs[SAMPLE_MAXB], newidx[SAMPLE_MAXB];
  __shared__ int tilesel[SAMPLE_MAXB];
  __shared__ double uu[SAMPLE_MAXB], ptt[SAMPLE_MAXB];
  __shared__ CollectLds cl;
  __shared__ int ntl;
  const WideHdr* h = ws.hdr;
  if (h->done) return;  // uniform (the state already says s = 0)
  const int tid = threadIdx.x;
  const int64_t i = h->i, N = h->N, ntiles = h->ntiles;
  const int s = h->s;
  double total = h->total;
  unsigned long long draws = h->draws;
  if (tid < h->r) newidx[tid] = ws.newidx[tid];
  __syncthreads();
  int nf = collect_first(newidx, h->r, found, 0, cl);
  draws += (unsigned long long)h->r;
  const uint64_t base[4] = {c0, c1, c2, c3};
  const WideLds L = wide_lds(wide_dyn, lds_ok, ntiles);
  // every round adds at least one new index, so s rounds suffice
  for (int round = 1; round <= s && nf < s; ++round) {
    const int n_f = nf;
    // sorted copy of the found positions, and the tiles that hold them
    if (tid == 0) {
      for (int q = 0; q < n_f; ++q) {
        int v = found[q], p = q;
        while (p > 0 && fs[p - 1] > v) {
          fs[p] = fs[p - 1];
          --p;
        }
        fs[p] = v;
      }
      int k = 0;
      for (int q = 0; q < n_f; ++q) {
        const int tt = fs[q] / WIDE_TILE;
        if (k == 0 || tilesel[k - 1] != tt) tilesel[k++] = tt;
      }
      ntl = k;
    }
    __syncthreads();
    const int nt = ntl;
    for (int q = 0; q < nt; ++q) {
      const int64_t tt = tilesel[q];
      double run[4];
      int c;
      wide_tile(diags, i, N, tt, fs, n_f, wsum, run, &c);
      if (tid == 1023) ws.S[tt] = run[3];
    }
    __syncthreads();  // (S read back by this workgroup only)
    total = wide_prefix(ws, ntiles, L, false, wsum, &bcast);
    const int r = s - n_f;
    if (tid < r) {
      const double u = rbrp_u64_to_unit(rbrp_philox_word(base, key0, key1, draws + tid));
      const int64_t t = wide_pick_tile(ws, L, ntiles, total, u);
      uu[tid] = u;
      newidx[tid] = (int)t;  // tile, resolved below
      ptt[tid] = wide_P(ws, L, t);
    }
    __syncthreads();
    for (int d = 0; d < r; ++d) {
      const int64_t j = wide_find(diags, i, N, ptt[d], newidx[d], total, uu[d], fs, n_f, wsum, imin);
      __syncthreads();
      if (tid == 0) newidx[d] = (int)j;
    }
    __syncthreads();
    nf = collect_first(newidx, r, found, n_f, cl);
    draws += (unsigned long long)r;
  }
  if (nf < s) {
    if (tid == 0) {
      bs->s = 0;
      bs->status = 4;
    }
    return;
  }
  for (int t = tid; t < s; t += 1024) J[t] = (int)(i + found[t]);
  if (tid == 0) {
    bs->s = s;
    bs->draws = draws;
    bs->nnz = h->nnz;
  }
}

// Bytes of LDS that k_wide_pick and k_wide_finish reserve for ntiles_max
// tiles on the current device: wide_lds_doubles(ntiles_max) doubles when that
// plus each kernel's static LDS fits the device's per-workgroup limit (and
// the allowance is granted), else 0 (the global-memory walks).
static size_t wide_lds_bytes(int64_t ntiles_max) {
  if (ntiles_max > WIDE_LDS_MAX_TILES) return 0;
  const size_t want = (size_t)wide_lds_doubles(ntiles_max) * sizeof(double);
  int dev = 0, maxlds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 0;
  hipFuncAttributes fp, ff;
  if (hipFuncGetAttributes(&fp, (const void*)k_wide_pick) != hipSuccess ||
      hipFuncGetAttributes(&ff, (const void*)k_wide_finish) != hipSuccess)
    return 0;
  const size_t stat = std::max(fp.sharedSizeBytes, ff.sharedSizeBytes);
  if (want + stat > (size_t)maxlds) return 0;
  static std::atomic<int> set_pick[RBRP_MAX_DEVICES], set_finish[RBRP_MAX_DEVICES];
  if (rbrp_allow_lds((const void*)k_wide_pick, (int)want, set_pick) != hipSuccess ||
      rbrp_allow_lds((const void*)k_wide_finish, (int)want, set_finish) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return want;
}

hipError_t rbrp_launch_sample(hipStream_t st, const double* diags, int64_t i, int64_t n,
                              int64_t kmax, int b, uint64_t key0, uint64_t key1,
                              const uint64_t ctr[4], double* w, double* cdf, RbrpBlockState* bs,
                              int* J, const int* i_dyn, double tol2) {
  if (b > SAMPLE_MAXB) return hipErrorInvalidValue;
  if (n - i - (i_dyn ? 1 : 0) > SAMPLE_WIDE_N) {
    // (with i_dyn the range is at most n - i - 1: the previous block kept a column)
    const int64_t ntiles_max = (n - i + WIDE_TILE - 1) / WIDE_TILE;
    const WideScratch ws = wide_scratch(w, cdf, ntiles_max);
    // up to four workgroups per compute unit, each looping over tiles
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    const int64_t tgrid = std::min<int64_t>(ntiles_max, 4 * (int64_t)cus);
    const size_t wl = wide_lds_bytes(ntiles_max);
    const int lds_ok = wl > 0 ? 1 : 0;
    hipLaunchKernelGGL(k_wide_tiles, dim3((unsigned)tgrid), dim3(1024), 0, st, diags, i, i_dyn, n, ws);
    hipLaunchKernelGGL(k_wide_pick, dim3(1), dim3(1024), wl, st, i, i_dyn, i_dyn ? 1 : 0, tol2, n,
                       kmax, b, key0, key1, ctr[0], ctr[1], ctr[2], ctr[3], lds_ok, ws, bs);
    hipLaunchKernelGGL(k_wide_resolve, dim3((unsigned)b), dim3(1024), 0, st, diags, ws);
    hipLaunchKernelGGL(k_wide_finish, dim3(1), dim3(1024), wl, st, diags, kmax, key0, key1, ctr[0],
                       ctr[1], ctr[2], ctr[3], lds_ok, ws, bs, J);
    return hipGetLastError();
  }
  // with i_dyn the first column is at least i + 1 (the previous block kept a
  // column, or the call does nothing)
  const bool in_lds = (n - i - (i_dyn ? 1 : 0)) <= SAMPLE_LDS_N;
  const size_t lds = in_lds ? 2 * SAMPLE_LDS_N * sizeof(double) : 0;
  if (lds > 48 * 1024) {
    static std::atomic<int> lds_set[RBRP_MAX_DEVICES];
    hipError_t e = rbrp_allow_lds((const void*)k_sample, (int)lds, lds_set);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_sample, dim3(1), dim3(1024), lds, st, diags, i, i_dyn, i_dyn ? 1 : 0, tol2,
                     in_lds ? 1 : 0, n, kmax, b, key0, key1, ctr[0], ctr[1], ctr[2], ctr[3], w, cdf,
                     bs, J);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pivoted Cholesky of the sampled block's Gram matrix G = V^T V (s x s) +
// robust filter.
//
// The pivot at step j is the first index of the largest trailing Schur
// diagonal, i.e. the column of V with the largest residual norm after
// projecting out the previous pivots: the pivot rule of LAPACK dgeqp3 (first
// maximum, with the swapped column moving to the vacated position), which the
// reference calls through magma_dgeqp3_gpu (gpu/rbrp.cu:410).  Row j of the
// factor is the row j of dgeqp3's R up to sign.
//
// Filter (gpu/rbrp.cu:193-207): sqr[r] = sum_{c>=r} R[r][c]^2, csum = running
// sum of sqr from the last row up, keep bt = s - k where k is the first index
// with csum[k] >= csum[s-1] / s (or / b for the nominal rule of
// RBRP.ipynb c11:69-73).
//
// Outputs: Jk = J[piv[0:bt]] (pivot order), Rb = R[0:bt][0:bt], its inverse,
// and the ascending symmetric differences Jin = sort(Jk) \ [i, i+bt),
// Jout = [i, i+bt) \ Jk (gpu/rbrp.cu:211-234).
//
// Blocks of at most 64 columns run the one-wave kernels (lane = column,
// k_panel.hip); larger ones the workgroup kernels with the Schur complement in
// registers.
// ---------------------------------------------------------------------------
hipError_t rbrp_launch_pivchol_filter(hipStream_t st, const double* G, int64_t ldg, int b,
                                      int filter_mode, int64_t i, const int* J, RbrpBlockState* bs,
                                      double* Rb, double* Rinv, int* Jk, int* Jin, int* Jout,
                                      int* Jkpos) {
  if (b <= 64)
    return rbrp_launch_pivchol_wave(st, G, ldg, b, filter_mode, i, J, bs, Rb, Rinv, Jk, Jin, Jout,
                                    Jkpos);
  return rbrp_launch_pivchol_regs(st, G, ldg, b, filter_mode, i, J, bs, Rb, Rinv, Jk, Jin, Jout,
                                  Jkpos);
}

// Second Cholesky-QR pass: G2 = Q1^T Q1 = R2^T R2.  Writes R2 (b x b, zeros
// below the diagonal), R2inv = R2^{-1} and Rout = R2 * Rb.
hipError_t rbrp_launch_chol2(hipStream_t st, const double* G2, int64_t ldg, int b,
                             RbrpBlockState* bs, const double* Rb, double* R2, double* R2inv,
                             double* Rout, int64_t ldr) {
  if (b <= 64) return rbrp_launch_chol2_wave(st, G2, ldg, b, bs, Rb, R2, R2inv, Rout, ldr);
  return rbrp_launch_chol2_regs(st, G2, ldg, b, bs, Rb, R2, R2inv, Rout, ldr);
}

// ---------------------------------------------------------------------------
// Symmetric-difference permutation (gpu/rbrp.cu:237-294), one launch:
//   perm: tmp = perm[Jk]; perm[Jin] = perm[Jout]; perm[i:i+bt] = tmp
//   diags: the same move as perm
//   R(0:i, :) columns: the same move as perm (matlab/RBRP.m:91-93)
//   Aw columns: Aw(:, Jin) = Aw(:, Jout) (the block columns then receive Q_b)
// The three parts touch disjoint data, so each workgroup takes one part:
//   workgroup 0: perm and diags (bt, nswap <= RBRP_MAX_B entries);
//   the next ceil(i / rt) workgroups: rt rows of R each, rt = PERM_R_VALS /
//     (2 b), every source value of those rows read into LDS before the
//     barrier and written to its destination after it (a row's sources and
//     destinations overlap: Jin lies inside Jk, Jout inside the front);
//   the rest: Aw in tiles of PERM_AW_ROWS rows x PERM_AW_COLS column pairs,
//     all loads of a thread before its stores (Jin, the kept columns outside
//     the front, and Jout, the front columns not kept, are disjoint).
// ---------------------------------------------------------------------------
#define PERM_R_VALS 4096
#define PERM_AW_ROWS 512
#define PERM_AW_COLS 8

static inline int64_t perm_r_rows(int b) { return PERM_R_VALS / (2 * b); }

template <typename T>
__global__ __launch_bounds__(256) void k_perm_block(int64_t i, const RbrpBlockState* bs,
                                                    const int* __restrict__ Jk,
                                                    const int* __restrict__ Jin,
                                                    const int* __restrict__ Jout, int* perm,
                                                    double* diags, double* R, int64_t ldr, int rt,
                                                    int r_groups, T* Aw, int64_t rows, int64_t lda,
                                                    int aw_tiles) {
  __shared__ double sv[PERM_R_VALS];
  __shared__ int src[2 * RBRP_MAX_B], dst[2 * RBRP_MAX_B];
  __shared__ double dk[RBRP_MAX_B], dout[RBRP_MAX_B];
  const int bt = bs->bt, ns = bs->nswap;
  const int tid = threadIdx.x;
  int g = blockIdx.x;
  if (g <= r_groups) {
    // source / destination columns of the moves: Jk -> the front, Jout -> Jin
    // (indices staged once, so that no data load waits on an index load)
    if (tid < bt + ns) {
      src[tid] = tid < bt ? Jk[tid] : Jout[tid - bt];
      dst[tid] = tid < bt ? (int)i + tid : Jin[tid - bt];
    }
    __syncthreads();
  }
  if (g == 0) {
    int pv = 0;
    if (tid < bt + ns) {
      pv = perm[src[tid]];
      if (tid < bt)
        dk[tid] = diags[src[tid]];
      else
        dout[tid - bt] = diags[src[tid]];
    }
    __syncthreads();
    if (tid >= bt && tid < bt + ns) {
      perm[dst[tid]] = pv;
      diags[dst[tid]] = dout[tid - bt];
    }
    __syncthreads();
    if (tid < bt) {
      perm[dst[tid]] = pv;
      diags[dst[tid]] = dk[tid];
    }
    return;
  }
  g -= 1;
  if (g < r_groups) {
    // slot s: source column s / rt (Jk, then Jout), row s % rt of the group;
    // all of a thread's loads are issued before the first LDS write waits
    const int64_t r0 = (int64_t)g * rt;
    const int nsl = rt * (bt + ns);
    double v[PERM_R_VALS / 256];
#pragma unroll
    for (int u = 0; u < PERM_R_VALS / 256; ++u) {
      const int s = tid + 256 * u;
      const int c = s / rt;
      const int64_t r = r0 + (s - c * rt);
      v[u] = 0.0;
      if (s < nsl && r < i) v[u] = R[r + (int64_t)src[c] * ldr];
    }
#pragma unroll
    for (int u = 0; u < PERM_R_VALS / 256; ++u)
      if (tid + 256 * u < nsl) sv[tid + 256 * u] = v[u];
    __syncthreads();
    for (int s = tid; s < nsl; s += 256) {
      const int c = s / rt;
      const int64_t r = r0 + (s - c * rt);
      if (r < i) R[r + (int64_t)dst[c] * ldr] = sv[s];
    }
    return;
  }
  g -= r_groups;
  if (g >= aw_tiles) return;
  const int rtiles = (int)((rows + PERM_AW_ROWS - 1) / PERM_AW_ROWS);
  const int cg = g / rtiles;
  const int c0 = cg * PERM_AW_COLS;
  if (c0 >= ns) return;
  const int nc = min(PERM_AW_COLS, ns - c0);
  // the tile's column indices, uniform over the workgroup (outside the
  // row-dependent branches, so they are read once before the data loads)
  int64_t co[PERM_AW_COLS], ci[PERM_AW_COLS];
#pragma unroll
  for (int c = 0; c < PERM_AW_COLS; ++c) {
    co[c] = c < nc ? (int64_t)Jout[c0 + c] * lda : 0;
    ci[c] = c < nc ? (int64_t)Jin[c0 + c] * lda : 0;
  }
  const int64_t rb = (int64_t)(g - cg * rtiles) * PERM_AW_ROWS + tid;
  T v[PERM_AW_COLS][PERM_AW_ROWS / 256];
#pragma unroll
  for (int c = 0; c < PERM_AW_COLS; ++c)
#pragma unroll
    for (int q = 0; q < PERM_AW_ROWS / 256; ++q) {
      const int64_t r = rb + 256 * q;
      if (c < nc && r < rows) v[c][q] = Aw[r + co[c]];
    }
#pragma unroll
  for (int c = 0; c < PERM_AW_COLS; ++c)
#pragma unroll
    for (int q = 0; q < PERM_AW_ROWS / 256; ++q) {
      const int64_t r = rb + 256 * q;
      if (c < nc && r < rows) Aw[r + ci[c]] = v[c][q];
    }
}

template <typename T>
static hipError_t launch_perm_block(hipStream_t st, int64_t i, int b, const RbrpBlockState* bs,
                                    const int* Jk, const int* Jin, const int* Jout, int* perm,
                                    double* diags, double* R, int64_t ldr, T* Aw, int64_t rows,
                                    int64_t lda) {
  if (b < 1 || b > RBRP_MAX_B || i < 0 || rows < 0) return hipErrorInvalidValue;
  const int64_t rt = perm_r_rows(b);
  const int64_t r_groups = (i + rt - 1) / rt;
  const int64_t aw_tiles = Aw ? ((rows + PERM_AW_ROWS - 1) / PERM_AW_ROWS) *
                                    ((b + PERM_AW_COLS - 1) / PERM_AW_COLS)
                              : 0;
  const int64_t grid = 1 + r_groups + aw_tiles;
  if (grid > INT_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_perm_block<T>, dim3((unsigned)grid), dim3(256), 0, st, i, bs, Jk, Jin,
                     Jout, perm, diags, R, ldr, (int)rt, (int)r_groups, Aw, rows, lda,
                     (int)aw_tiles);
  return hipGetLastError();
}

hipError_t rbrp_launch_perm_block(hipStream_t st, int64_t i, int b, const RbrpBlockState* bs,
                                  const int* Jk, const int* Jin, const int* Jout, int* perm,
                                  double* diags, double* R, int64_t ldr, double* Aw, int64_t rows,
                                  int64_t lda) {
  return launch_perm_block<double>(st, i, b, bs, Jk, Jin, Jout, perm, diags, R, ldr, Aw, rows, lda);
}

hipError_t rbrp_launch_perm_block32(hipStream_t st, int64_t i, int b, const RbrpBlockState* bs,
                                    const int* Jk, const int* Jin, const int* Jout, int* perm,
                                    double* diags, double* R, int64_t ldr, float* Aw, int64_t rows,
                                    int64_t lda) {
  return launch_perm_block<float>(st, i, b, bs, Jk, Jin, Jout, perm, diags, R, ldr, Aw, rows, lda);
}

// diags[i + bt + c] = sum_{t < ntiles} part[t * ldp + c], c < n - i - bt.
// Residual total over diags[c0, n) by one 1024-thread workgroup (a fixed
// summation order), stored in bs and, with `pub`, published to the host: the
// fields are written first, then `seq` with a system-scope release, so a host
// that reads seq == the expected number (with acquire) sees the whole state.
__device__ void resid_total_publish(const double* __restrict__ diags, int64_t c0, int64_t n,
                                    RbrpBlockState* bs, RbrpBlockState* pub, RbrpBlockState* snap,
                                    int seq, double* red) {
  double acc = 0.0;
  for (int64_t j = c0 + threadIdx.x; j < n; j += 1024) acc += diags[j];
  const double t = rbrp_block_sum(acc, red);
  if (threadIdx.x == 0) {
    const double err = (bs->N0 > 0.0) ? t / bs->N0 : 0.0;
    bs->resid = t;
    bs->err = err;
    RbrpBlockState rec;
    rec.s = bs->s;
    rec.bt = bs->bt;
    rec.nswap = bs->nswap;
    rec.status = bs->status;
    rec.draws = bs->draws;
    rec.resid = t;
    rec.err = err;
    rec.N0 = bs->N0;
    rec.nnz = bs->nnz;
    rec.seq = seq;
    // device snapshot of this block's record: no later kernel writes it, so
    // the host can copy it after the stream has gone idle
    if (snap) *snap = rec;
    if (pub) {
      pub->s = rec.s;
      pub->bt = rec.bt;
      pub->nswap = rec.nswap;
      pub->status = rec.status;
      pub->draws = rec.draws;
      pub->resid = rec.resid;
      pub->err = rec.err;
      pub->N0 = rec.N0;
      pub->nnz = rec.nnz;
      __hip_atomic_store(&pub->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// Squared norms of 64 columns from the per-row-group partials: wave w sums
// tiles w, w+16, ... (lane = column, so every load is a 512-byte run), then the
// 16 wave partials are added in wave order.
__device__ void colnorm_tile(const double* __restrict__ part, int64_t ldp, int64_t ntiles,
                             int64_t c0, int64_t c, int64_t n, double* __restrict__ diags,
                             double (*red)[64]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (c0 + c < n) {
    const double* p = part + c;
    int64_t t = wid;
    for (; t + 48 < ntiles; t += 64) {
      a0 += p[t * ldp];
      a1 += p[(t + 16) * ldp];
      a2 += p[(t + 32) * ldp];
      a3 += p[(t + 48) * ldp];
    }
    for (; t < ntiles; t += 16) a0 += p[t * ldp];
  }
  red[wid][lane] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  if (wid == 0 && c0 + c < n) {
    double acc = 0.0;
    for (int w = 0; w < 16; ++w) acc += red[w][lane];
    diags[c0 + c] = acc;
  }
  __syncthreads();  // red is reused by the caller
}

// One workgroup per 64 columns (colnorm_tile).
__global__ __launch_bounds__(1024) void k_colnorm_reduce(const double* __restrict__ part,
                                                         int64_t ldp, int64_t ntiles, int64_t i,
                                                         int64_t n, const RbrpBlockState* bs,
                                                         double* __restrict__ diags) {
  __shared__ double red[16][64];
  const int64_t c0 = i + bs->bt;
  const int64_t c = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  if (c0 + (int64_t)blockIdx.x * 64 < n) colnorm_tile(part, ldp, ntiles, c0, c, n, diags, red);
}

hipError_t rbrp_launch_colnorm_reduce(hipStream_t st, const double* part, int64_t ldp, int64_t ntiles,
                                      int64_t i, int64_t n, RbrpBlockState* bs, double* diags) {
  const int64_t cmax = n - i;
  if (cmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_colnorm_reduce, dim3((unsigned)((cmax + 63) / 64)), dim3(1024), 0, st, part,
                     ldp, ntiles, i, n, bs, diags);
  return hipGetLastError();
}

// resid = sum(diags[i + bt : n]); err = resid / N0 (gpu/rbrp.cu:467).
// Residual total after the block (see resid_total_publish).
__global__ __launch_bounds__(1024) void k_resid_total(const double* __restrict__ diags, int64_t i,
                                                      int64_t n, RbrpBlockState* bs,
                                                      RbrpBlockState* pub, RbrpBlockState* snap,
                                                      int seq) {
  __shared__ double red[16];
  resid_total_publish(diags, i + bs->bt, n, bs, pub, snap, seq, red);
}

hipError_t rbrp_launch_resid_total(hipStream_t st, const double* diags, int64_t i, int64_t n,
                                   RbrpBlockState* bs, RbrpBlockState* pub, int seq,
                                   RbrpBlockState* snap) {
  hipLaunchKernelGGL(k_resid_total, dim3(1), dim3(1024), 0, st, diags, i, n, bs, pub, snap, seq);
  return hipGetLastError();
}

// dst(0:rows, 0:cols) = src(0:rows, 0:cols)
__global__ void k_copy_block(int64_t rows, const double* __restrict__ src, int64_t lds,
                             double* __restrict__ dst, int64_t ldd) {
  const int64_t r = (int64_t)blockIdx.y * 256 + threadIdx.x;
  const int64_t c = blockIdx.x;
  if (r >= rows) return;
  dst[r + c * ldd] = src[r + c * lds];
}

hipError_t rbrp_launch_copy_block(hipStream_t st, int64_t rows, int64_t cols, const double* src,
                                  int64_t lds, double* dst, int64_t ldd) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_block, dim3((unsigned)cols, (unsigned)((rows + 255) / 256)), dim3(256), 0,
                     st, rows, src, lds, dst, ldd);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Diagonal-block back substitution of X = R11^{-1} R12: X(0:nb, c) =
// Rd^{-1} X(0:nb, c) for every column c; Rd is nb x nb upper (nb <= 64),
// padded to a 64 x 64 unit triangle.  One thread per column, its 64 entries in
// registers (fully unrolled substitution); Rd in LDS, read as broadcasts.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_trsm_diag(int nb, int64_t ncols, const double* __restrict__ Rd,
                                                  int64_t ldr, double* __restrict__ X, int64_t ldx) {
  __shared__ double Rs[64 * 65];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * 64; idx += 64) {
    const int r = idx & 63, c = idx >> 6;
    Rs[r * 65 + c] = (r < nb && c < nb) ? Rd[r + (int64_t)c * ldr] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * 64 + tid;
  if (c >= ncols) return;
  double* xc = X + c * ldx;
  double x[64];
#pragma unroll
  for (int r = 0; r < 64; ++r) x[r] = (r < nb) ? xc[r] : 0.0;
#pragma unroll
  for (int r = 63; r >= 0; --r) {
    double v = x[r];
#pragma unroll
    for (int l = r + 1; l < 64; ++l) v -= Rs[r * 65 + l] * x[l];
    x[r] = v / Rs[r * 65 + r];
  }
#pragma unroll
  for (int r = 0; r < 64; ++r)
    if (r < nb) xc[r] = x[r];
}

hipError_t rbrp_launch_trsm_diag(hipStream_t st, int64_t nb, int64_t ncols, const double* Rdiag,
                                 int64_t ldr, double* X, int64_t ldx) {
  if (nb <= 0 || ncols <= 0) return hipSuccess;
  if (nb > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_trsm_diag, dim3((unsigned)((ncols + 63) / 64)), dim3(64), 0, st, (int)nb, ncols,
                     Rdiag, ldr, X, ldx);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// X(0:k, c) <- R^{-1} X(0:k, c) for the k x k upper-triangular R (ldr) in one
// launch, one workgroup per 16 columns of X: the 64-row blocks from the bottom
// (block q = rows [max(0, k - 64 (q+1)), k - 64 q), the partition of
// k_upper_inv_blocks, whose inverses are at dinv + 4096 q, ld 64), each as
//   T = X(blk, c) - R(blk, below) X(below, c),   X(blk, c) = Rinv(blk) T
// on the FP64 matrix core.  Sixteen waves: wave (rw, ks) takes rows 16 rw ..
// 16 rw + 15 of the block and every TRSM_KS-th 16-deep step of the rows below
// from step ks (the 4 q steps of block q split evenly); MFMA s of a step takes
// row kb + 4 g + s of X from lane group g, and the slice's next step is loaded
// before the current step's products (clamped to the last step, so every load
// is unconditional).  The slices' partial sums meet in LDS and are added in
// slice order; the first slice then forms T, multiplies it by the block
// inverse (entries past the block's size masked) and stores the block's rows,
// a workgroup barrier ordering them before the next block's loads.
// ---------------------------------------------------------------------------
#define TRSM_KS 4
__global__ __launch_bounds__(256 * TRSM_KS) void k_trsm_cols(int64_t k, int64_t nc,
                                                             const double* __restrict__ R, int64_t ldr,
                                                             const double* __restrict__ dinv,
                                                             double* X, int64_t ldx) {
  __shared__ double Tp[TRSM_KS][64 * 17];  // slice partials, then T in Tp[0]: (t, col) at t * 17 + col
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int rw = w & 3, ks = w >> 2;
  const int j = lane & 15, g = lane >> 4;
  const int64_t c0 = (int64_t)blockIdx.x * 16;
  if (c0 >= nc) return;  // uniform per workgroup
  const bool cok = c0 + j < nc;
  double* xc = X + min(c0 + j, nc - 1) * ldx;  // this lane's column (clamped)
  const int nq = (int)((k + 63) / 64);
  for (int q = 0; q < nq; ++q) {
    const int64_t r1 = k - 64 * (int64_t)q, r0 = max((int64_t)0, r1 - 64);
    const int nb = (int)(r1 - r0);
    // A operand: row r0 + 16 rw + j of R (clamped into the block)
    const double* rrow = R + r0 + min(16 * rw + j, nb - 1);
    v4d acc = (v4d){0.0, 0.0, 0.0, 0.0};
    if (q > 0) {
      // this slice's steps: r1 + 16 (ks + TRSM_KS t), t < q (k - r1 = 64 q)
      const int64_t klast = k - 16;
      int64_t kb = r1 + 16 * ks;
      double a[4], b[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a[s] = rrow[(kb + 4 * g + s) * ldr];
        b[s] = xc[kb + 4 * g + s];
      }
      for (int t = 0; t < q; ++t, kb += 16 * TRSM_KS) {
        const int64_t kn = min(kb + 16 * TRSM_KS, klast);
        double an[4], bn[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          an[s] = rrow[(kn + 4 * g + s) * ldr];
          bn[s] = xc[kn + 4 * g + s];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          a[s] = an[s];
          b[s] = bn[s];
        }
      }
    }
    // lane (j, g) holds row 16 rw + 4 v + g of the slice's product in acc[v]
#pragma unroll
    for (int v = 0; v < 4; ++v) Tp[ks][(16 * rw + 4 * v + g) * 17 + j] = acc[v];
    __syncthreads();
    if (ks == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rl = 16 * rw + 4 * v + g;
        double sum = Tp[0][rl * 17 + j];
#pragma unroll
        for (int z = 1; z < TRSM_KS; ++z) sum += Tp[z][rl * 17 + j];
        Tp[0][rl * 17 + j] = (rl < nb) ? xc[r0 + rl] - sum : 0.0;
      }
    }
    __syncthreads();
    if (ks == 0) {
      const double* dq = dinv + 4096 * (int64_t)q;  // Rinv(row, t) at dq[row + 64 t]
      const int row = 16 * rw + j;
      v4d y = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int tb = 0; tb < 64; tb += 16)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int t = tb + 4 * g + s;
          const double av = (row < nb && t < nb) ? dq[row + 64 * t] : 0.0;
          y = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Tp[0][t * 17 + j], y, 0, 0, 0);
        }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rl = 16 * rw + 4 * v + g;
        if (rl < nb && cok) xc[r0 + rl] = y[v];
      }
    }
    __syncthreads();  // the block's rows are stored and Tp is free before the next block
  }
}

hipError_t rbrp_launch_trsm_cols(hipStream_t st, int64_t k, int64_t nc, const double* R, int64_t ldr,
                                 const double* dinv, double* X, int64_t ldx) {
  if (k <= 0 || nc <= 0) return hipSuccess;
  if (!dinv || ldr < k || ldx < k) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_trsm_cols, dim3((unsigned)((nc + 15) / 16)), dim3(256 * TRSM_KS), 0, st, k, nc,
                     R, ldr, dinv, X, ldx);
  return hipGetLastError();
}

// out[t] = uniform number draw0 + t of the stream (key0, key1, ctr).
__global__ void k_uniforms(uint64_t key0, uint64_t key1, uint64_t c0, uint64_t c1, uint64_t c2,
                           uint64_t c3, uint64_t draw0, int64_t count, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  const uint64_t base[4] = {c0, c1, c2, c3};
  out[t] = rbrp_u64_to_unit(rbrp_philox_word(base, key0, key1, draw0 + (uint64_t)t));
}

hipError_t rbrp_launch_uniforms(hipStream_t st, uint64_t key0, uint64_t key1, const uint64_t ctr[4],
                                uint64_t draw0, int64_t count, double* out) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_uniforms, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, key0, key1,
                     ctr[0], ctr[1], ctr[2], ctr[3], draw0, count, out);
  return hipGetLastError();
}
