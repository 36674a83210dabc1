// Batched RBRP column IDs in one persistent launch (gfx950).
//
// Each workgroup (4 waves) repeatedly takes the next problem from a device
// work-queue counter and runs the whole RBRP loop of that problem in its own
// workspace slot: copy + column norms, then per block sampling, second
// Gram-Schmidt pass, Gram + pivoted Cholesky + robust filter, CholQR2,
// symmetric-difference permutation, projection and update with exact
// residual norms, and at the end X = R11^{-1} R12.  The steps are the ones of
// the single-problem loop (rbrp_host.cpp), so a problem gives the same
// skeleton whichever path runs it; all of a problem's arithmetic stays in one
// workgroup (no launches, no host round trips), and thousands of uneven
// problems share the chip through the queue.
//
// Tall-by-small products use v_mfma_f64_16x16x4f64 on 16 x 16 tiles inside
// the workgroup (operand layout: A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, D[i][j] in lane j + 16*(i%4), element i/4).  Small dense work (the
// s x s panel, at most 32 x 32) runs in one wave out of LDS.
#include <algorithm>

#include "rbrp_batched.h"
#include "rbrp_kernels.h"

#define BQ_NT 256

// ---------------------------------------------------------------------------
// LDS carve-up (dynamic, sized from n_max and b at launch)
// ---------------------------------------------------------------------------
struct BqLds {
  double *diags, *w, *cdf;                       // n_max each
  double *G, *Rb, *Rinv, *G2, *R2, *R2inv, *Rn;  // b x (b + 1) each, [r * ldS + c]
  double *part;                                  // 4 x 256 (K-split partial tiles)
  double *dk, *dout;                             // b each
  double *red;                                   // 8
  int *J, *Jk, *Jkpos, *Jin, *Jout, *found, *newidx, *isnew, *tk, *to;  // b each
  int *ired;                                     // 8
  int *sc;                                       // scalars
  int ldS;
};

enum { SC_S = 0, SC_BT, SC_NSWAP, SC_STATUS, SC_NFOUND, SC_PROB, SC_RANK, SC_COUNT };

__host__ __device__ inline size_t bq_lds_bytes(int n_max, int b) {
  const size_t ldS = (size_t)b + 1;
  size_t d = 3 * (size_t)n_max + 7 * (size_t)b * ldS + 4 * 256 + 2 * (size_t)b + 8;
  size_t i = 10 * (size_t)b + 8 + SC_COUNT;
  return d * sizeof(double) + i * sizeof(int);
}

__device__ inline BqLds bq_carve(double* base, int n_max, int b) {
  BqLds L;
  L.ldS = b + 1;
  const int sz = b * L.ldS;
  double* p = base;
  L.diags = p; p += n_max;
  L.w = p; p += n_max;
  L.cdf = p; p += n_max;
  L.G = p; p += sz;
  L.Rb = p; p += sz;
  L.Rinv = p; p += sz;
  L.G2 = p; p += sz;
  L.R2 = p; p += sz;
  L.R2inv = p; p += sz;
  L.Rn = p; p += sz;
  L.part = p; p += 4 * 256;
  L.dk = p; p += b;
  L.dout = p; p += b;
  L.red = p; p += 8;
  int* q = (int*)p;
  L.J = q; q += b;
  L.Jk = q; q += b;
  L.Jkpos = q; q += b;
  L.Jin = q; q += b;
  L.Jout = q; q += b;
  L.found = q; q += b;
  L.newidx = q; q += b;
  L.isnew = q; q += b;
  L.tk = q; q += b;
  L.to = q; q += b;
  L.ired = q; q += 8;
  L.sc = q;
  return L;
}

// ---------------------------------------------------------------------------
// Workgroup MFMA helpers
// ---------------------------------------------------------------------------
// Out(r, c) = sum_{k < K} A(k, r) B(k, c) for r < M, c < N.  colA(r) / colB(c)
// return the FP64 column pointers; K is a multiple of 16 and both operands
// are zero in rows [true height, K).  With fewer than four 16 x 16 output
// tiles the K range is split over the waves and the partial tiles are added
// in wave order (deterministic).  store(r, c, v) receives each entry.
template <class FA, class FB, class FS>
__device__ void bq_tn(int M, int N, int K, FA colA, FB colB, double* part, FS store) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int TM = (M + 15) >> 4, TNn = (N + 15) >> 4, T = TM * TNn;
  if (T == 0) return;  // uniform
  const int nsplit = T >= 4 ? 1 : 4 / T;
  const int kc = (((K + nsplit - 1) / nsplit) + 15) & ~15;
  for (int item = w; item < T * nsplit; item += 4) {
    const int tile = item / nsplit, pid = item % nsplit;
    const int r0 = (tile % TM) * 16, c0 = (tile / TM) * 16;
    const double* pa = colA(min(r0 + li, M - 1));
    const double* pb = colB(min(c0 + li, N - 1));
    const int kb = pid * kc, ke = min(K, kb + kc);
    // four 16-row steps per iteration: sixteen loads in flight per lane, the
    // MFMAs alternating over two accumulators
    v4d acc = {0.0, 0.0, 0.0, 0.0}, acc2 = acc;
    int k0 = kb;
    for (; k0 + 64 <= ke; k0 += 64) {
      double2 a[8], bv[8];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[2 * u] = *(const double2*)(pa + k0 + 16 * u + 4 * lg);
        a[2 * u + 1] = *(const double2*)(pa + k0 + 16 * u + 4 * lg + 2);
        bv[2 * u] = *(const double2*)(pb + k0 + 16 * u + 4 * lg);
        bv[2 * u + 1] = *(const double2*)(pb + k0 + 16 * u + 4 * lg + 2);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x, bv[u].x, acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].y, bv[u].y, acc2, 0, 0, 0);
      }
    }
    for (; k0 < ke; k0 += 16) {
      const double2 a0 = *(const double2*)(pa + k0 + 4 * lg);
      const double2 a1 = *(const double2*)(pa + k0 + 4 * lg + 2);
      const double2 b0 = *(const double2*)(pb + k0 + 4 * lg);
      const double2 b1 = *(const double2*)(pb + k0 + 4 * lg + 2);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.x, b0.x, acc, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, b0.y, acc2, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.x, b1.x, acc, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, b1.y, acc2, 0, 0, 0);
    }
    acc += acc2;
    if (nsplit == 1) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + 4 * v + lg, c = c0 + li;
        if (r < M && c < N) store(r, c, acc[v]);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) part[w * 256 + lane * 4 + v] = acc[v];
    }
  }
  if (nsplit > 1) {
    __syncthreads();
    if (w < T * nsplit && w % nsplit == 0) {
      const int tile = w / nsplit;
      const int r0 = (tile % TM) * 16, c0 = (tile / TM) * 16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double s = 0.0;
        for (int q = 0; q < nsplit; ++q) s += part[(w + q) * 256 + lane * 4 + v];
        const int r = r0 + 4 * v + lg, c = c0 + li;
        if (r < M && c < N) store(r, c, s);
      }
    }
  }
  __syncthreads();
}

// C(r, c) = [C(r, c) -] sum_{k < K} A(r, k) B(k, c) for r < rows, c < N.  The
// MFMA computes the transposed block, so lane l covers row r0 + l%16 and
// columns c0 + 4v + l/16: every C access moves four 128-byte runs.  With four
// or more 16-column tiles wave w owns the column tiles w, w + 4, ...; with
// fewer, each column tile is dealt to 4 / tiles waves that take every
// (4 / tiles)-th 16-row tile, so no wave idles on the narrow products (the
// re-orthogonalisation, CholQR2 and late updates).  A wave takes BQ_RT of its
// row tiles at a time with every load of those issued before their MFMAs
// (BQ_RT > 1 only where the register budget of the kernel allows it).
// With NORMS, normout(c, sum over rows of the new C(r, c)^2) for each column;
// the partial sums of the waves sharing a column tile go through part (4 x 16
// doubles) and are added in wave order.  colA(k) / colC(c) give column
// pointers (rows addressed directly); belem(k, c) gives B.  Rows >= rows are
// neither read nor written.
template <bool SUB, bool NORMS, int BQ_RT, class FA, class FB, class FC, class FN>
__device__ void bq_nn(int rows, int N, int K, FA colA, FB belem, FC colC, FN normout,
                      double* part) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int TNn = (N + 15) >> 4, nrt = (rows + 15) >> 4;
  const int ws = TNn >= 4 ? 1 : 4 / max(TNn, 1);  // waves per column tile (uniform)
  const int sub = w % ws;
  for (int tn = w / ws; tn < TNn; tn += 4 / ws) {
    const int c0 = tn * 16;
    const bool cok = c0 + li < N;
    double nsq[4] = {0.0, 0.0, 0.0, 0.0};
    // B fragment of the first 16 k (the whole reduction for K <= 16: the
    // update and the CholQR2 products), loaded once for the column tile
    double bq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = 4 * q + lg;
      bq[q] = (k < K && cok) ? belem(k, c0 + li) : 0.0;
    }
    for (int t0 = sub; t0 < nrt; t0 += BQ_RT * ws) {
      // row tiles t0 + u * ws, u < BQ_RT
      int rr[BQ_RT];
      bool rok[BQ_RT];
#pragma unroll
      for (int u = 0; u < BQ_RT; ++u) {
        rr[u] = (t0 + u * ws) * 16 + li;
        rok[u] = rr[u] < rows;
      }
      double cv[BQ_RT][4];
      if (SUB) {
#pragma unroll
        for (int u = 0; u < BQ_RT; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            cv[u][v] = rok[u] ? colC(min(c0 + 4 * v + lg, N - 1))[rr[u]] : 0.0;
      }
      v4d acc[BQ_RT];
#pragma unroll
      for (int u = 0; u < BQ_RT; ++u) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < K; k0 += 16) {
        double av[BQ_RT][4], bv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = k0 + 4 * q + lg;
          bv[q] = k0 == 0 ? bq[q] : ((k < K && cok) ? belem(k, c0 + li) : 0.0);
#pragma unroll
          for (int u = 0; u < BQ_RT; ++u) av[u][q] = (k < K && rok[u]) ? colA(k)[rr[u]] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int u = 0; u < BQ_RT; ++u)
            acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[q], av[u][q], acc[u], 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < BQ_RT; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = c0 + 4 * v + lg;
          if (c < N && rok[u]) {
            const double nv = SUB ? cv[u][v] - acc[u][v] : acc[u][v];
            colC(c)[rr[u]] = nv;
            nsq[v] = fma(nv, nv, nsq[v]);
          }
        }
    }
    if (NORMS) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double x = nsq[v];
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        x += __shfl_xor(x, 8, 64);
        const int c = c0 + 4 * v + lg;
        if (ws == 1) {
          if (li == 0 && c < N) normout(c, x);
        } else if (li == 0) {
          part[w * 16 + 4 * v + lg] = x;
        }
      }
    }
  }
  if (NORMS && ws > 1) {
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 16 * TNn) {
      const int tn = t >> 4, cl = t & 15;
      double s = 0.0;
      for (int q = 0; q < ws; ++q) s += part[(tn * ws + q) * 16 + cl];
      if (tn * 16 + cl < N) normout(tn * 16 + cl, s);
    }
  }
  __syncthreads();
}

// Projection and update of the trailing columns fused per 16-column tile, for
// bt <= 16 kept columns and N >= 49 trailing columns (every wave owns a
// tile): wave w takes the tiles w, w + 4, ...; for each it forms
// W = Q^T C(:, tile) over all K rows (the loop of bq_tn), stores W into
// Wout (row r = kept column, leading dimension ldw), and at once updates the
// same columns C(:, tile) -= Q W with their new squared norms (the loop of
// bq_nn with BQ_RT = RT row tiles a wave at a time, RT a launch shape the
// tuner picks: every element's sum is the same for any RT): the MFMA output
// of the projection, lane (j, g)
// holding W(4v + g, c0 + j), is already the B fragment of the update.  The
// arithmetic is that of bq_tn followed by bq_nn, without the barrier between
// them, and the second pass over a tile follows its first while its rows
// are still in the caches.  Q (K x bt) and C (K x N) have leading dimension K
// (a multiple of 16, rows >= rows zero in Q); rows >= rows of C are neither
// read nor written by the update.
template <int RT, class FN>
__device__ void bq_proj_update(int rows, int K, int bt, int N, const double* Q, double* C,
                               double* Wout, int64_t ldw, FN normout) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int TNn = (N + 15) >> 4, nrt = (rows + 15) >> 4;
  const double* pa = Q + (int64_t)min(li, bt - 1) * K;
  for (int tn = w; tn < TNn; tn += 4) {
    const int c0 = tn * 16;
    const bool cok = c0 + li < N;
    const double* pb = C + (int64_t)min(c0 + li, N - 1) * K;
    // ---- W(r, c0 + li) for r = 4v + lg
    v4d acc = {0.0, 0.0, 0.0, 0.0}, acc2 = acc;
    int k0 = 0;
    for (; k0 + 64 <= K; k0 += 64) {
      double2 a[8], bv[8];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[2 * u] = *(const double2*)(pa + k0 + 16 * u + 4 * lg);
        a[2 * u + 1] = *(const double2*)(pa + k0 + 16 * u + 4 * lg + 2);
        bv[2 * u] = *(const double2*)(pb + k0 + 16 * u + 4 * lg);
        bv[2 * u + 1] = *(const double2*)(pb + k0 + 16 * u + 4 * lg + 2);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x, bv[u].x, acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].y, bv[u].y, acc2, 0, 0, 0);
      }
    }
    for (; k0 < K; k0 += 16) {
      const double2 a0 = *(const double2*)(pa + k0 + 4 * lg);
      const double2 a1 = *(const double2*)(pa + k0 + 4 * lg + 2);
      const double2 b0 = *(const double2*)(pb + k0 + 4 * lg);
      const double2 b1 = *(const double2*)(pb + k0 + 4 * lg + 2);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.x, b0.x, acc, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, b0.y, acc2, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.x, b1.x, acc, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, b1.y, acc2, 0, 0, 0);
    }
    acc += acc2;
    // rows r >= bt of acc come from the clamped column of Q: zero them for the
    // update; the kept rows go to Wout
    double bq[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 4 * v + lg;
      bq[v] = (r < bt && cok) ? acc[v] : 0.0;
      if (r < bt && cok) Wout[r + (int64_t)(c0 + li) * ldw] = acc[v];
    }
    // ---- C(:, tile) -= Q W(:, tile), with the new squared norms
    double nsq[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < nrt; t0 += RT) {
      int rr[RT];
      bool rok[RT];
#pragma unroll
      for (int u = 0; u < RT; ++u) {
        rr[u] = (t0 + u) * 16 + li;
        rok[u] = rr[u] < rows;
      }
      double cv[RT][4], av[RT][4];
#pragma unroll
      for (int u = 0; u < RT; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          cv[u][v] = rok[u] ? C[rr[u] + (int64_t)min(c0 + 4 * v + lg, N - 1) * K] : 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = 4 * q + lg;
#pragma unroll
        for (int u = 0; u < RT; ++u) av[u][q] = (k < bt && rok[u]) ? Q[rr[u] + (int64_t)k * K] : 0.0;
      }
      v4d up[RT];
#pragma unroll
      for (int u = 0; u < RT; ++u) up[u] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int u = 0; u < RT; ++u)
          up[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(bq[q], av[u][q], up[u], 0, 0, 0);
#pragma unroll
      for (int u = 0; u < RT; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = c0 + 4 * v + lg;
          if (c < N && rok[u]) {
            const double nv = cv[u][v] - up[u][v];
            C[rr[u] + (int64_t)c * K] = nv;
            nsq[v] = fma(nv, nv, nsq[v]);
          }
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double x = nsq[v];
      x += __shfl_xor(x, 1, 64);
      x += __shfl_xor(x, 2, 64);
      x += __shfl_xor(x, 4, 64);
      x += __shfl_xor(x, 8, 64);
      const int c = c0 + 4 * v + lg;
      if (li == 0 && c < N) normout(c, x);
    }
  }
  __syncthreads();
}

// *dst(idx) = *src(idx) for idx < count over the workgroup's threads, the
// loads of four iterations issued before their stores (a copy otherwise
// waits for each load in turn).  Source and destination must not overlap.
template <class FS, class FD>
__device__ inline void bq_copy(int count, FS src, FD dst) {
  int idx = threadIdx.x;
  for (; idx + 3 * BQ_NT < count; idx += 4 * BQ_NT) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = *src(idx + u * BQ_NT);
#pragma unroll
    for (int u = 0; u < 4; ++u) *dst(idx + u * BQ_NT) = v[u];
  }
  for (; idx < count; idx += BQ_NT) *dst(idx) = *src(idx);
}

// ---------------------------------------------------------------------------
// Weighted sampling without replacement: NumPy's Generator.choice rounds on
// the Philox stream (key (k0, 0), counter base); the same algorithm as
// k_sample, for the 256 threads of one workgroup and weights in LDS.
// Returns s (uniform); sets sc[SC_STATUS] = 4 if a round fails to progress.
// ---------------------------------------------------------------------------
__device__ int bq_sample(const BqLds& L, int i, int n, int kmax, int b, uint64_t k0,
                         const uint64_t base[4], unsigned long long& draws) {
  const int tid = threadIdx.x;
  const int N = n - i;
  int cnt = 0;
  for (int j = tid; j < N; j += BQ_NT) {
    double v = L.diags[i + j];
    v = (v > 0.0) ? v : 0.0;
    L.w[j] = v;
    cnt += (v > 0.0) ? 1 : 0;
  }
  const int nnz = rbrp_block_sum_int(cnt, L.ired);
  int s = b;
  if (kmax - i < s) s = kmax - i;
  if (nnz < s) s = nnz;
  if (s <= 0) return 0;
  if (tid == 0) L.sc[SC_NFOUND] = 0;
  __syncthreads();
  const int chunk = (N + BQ_NT - 1) / BQ_NT;
  const int j0 = min(N, tid * chunk), j1 = min(N, j0 + chunk);
  for (int round = 0; round <= s; ++round) {
    const int nf = L.sc[SC_NFOUND];
    if (nf >= s) break;  // uniform
    const int r = s - nf;
    for (int t = tid; t < nf; t += BQ_NT) L.w[L.found[t]] = 0.0;
    __syncthreads();
    double loc = 0.0;
    for (int j = j0; j < j1; ++j) loc += L.w[j];
    // exclusive scan of the chunk sums over the workgroup (wave scan + wave offsets)
    const int lane = tid & 63, wid = tid >> 6;
    double x = loc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    double ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = 0.0;
    if (lane == 63) L.red[wid] = x;
    __syncthreads();
    double run = ex;
    for (int q = 0; q < wid; ++q) run += L.red[q];
    for (int j = j0; j < j1; ++j) {
      run += L.w[j];
      L.cdf[j] = run;
    }
    __syncthreads();
    const double total = L.cdf[N - 1];
    if (tid < r) {
      const double u = rbrp_u64_to_unit(rbrp_philox_word(base, k0, 0, draws + tid));
      int lo = 0, hi = N;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L.cdf[mid] / total > u)
          hi = mid;
        else
          lo = mid + 1;
      }
      if (lo >= N) lo = N - 1;
      L.newidx[tid] = lo;
    }
    __syncthreads();
    if (tid < r) {
      const int v = L.newidx[tid];
      int dup = 0;
      for (int q = 0; q < nf; ++q) dup |= (L.found[q] == v) ? 1 : 0;
      for (int q = 0; q < tid; ++q) dup |= (L.newidx[q] == v) ? 1 : 0;
      L.isnew[tid] = !dup;
    }
    __syncthreads();
    if (tid < r && L.isnew[tid]) {
      int pos = nf;
      for (int q = 0; q < tid; ++q) pos += L.isnew[q];
      L.found[pos] = L.newidx[tid];
    }
    __syncthreads();
    if (tid == 0) {
      int add = 0;
      for (int q = 0; q < r; ++q) add += L.isnew[q];
      L.sc[SC_NFOUND] = nf + add;
    }
    draws += (unsigned long long)r;
    __syncthreads();
  }
  if (L.sc[SC_NFOUND] < s) {
    if (tid == 0) L.sc[SC_STATUS] = 4;
    __syncthreads();
    return 0;
  }
  for (int t = tid; t < s; t += BQ_NT) L.J[t] = i + L.found[t];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// One-wave panel routines on LDS matrices [r * ldS + c] of at most 32 x 32.
// ---------------------------------------------------------------------------
// X = U^{-1} (n x n upper triangular, U, X with pitch ld; X's lower part 0).
__device__ inline void bq_upper_inverse(const double* U, double* X, int n, int ld) {
  const int t = threadIdx.x;  // lane t: column t
  if (t < 64) {
    for (int r = 0; r < n; ++r)
      if (t < n) X[r * ld + t] = 0.0;
    if (t < n) X[t * ld + t] = 1.0 / U[t * ld + t];
    for (int r = n - 2; r >= 0; --r) {
      if (t < n && t > r) {
        double acc = 0.0;
        for (int l = r + 1; l <= t; ++l) acc = fma(U[r * ld + l], X[l * ld + t], acc);
        X[r * ld + t] = -acc / U[r * ld + r];
      }
    }
  }
}

// Pivoted Cholesky of the s x s Gram matrix G (dgeqp3 pivot rule: largest
// Schur diagonal, lowest position on ties; the column at the pivot's position
// takes the vacated one), robust filter, kept factor Rb and its inverse, and
// the symmetric-difference lists.  Wave 0 works; every thread must call.
__device__ void bq_pivchol(const BqLds& L, int s, int b, int filter_mode, int i) {
  const int c = threadIdx.x;
  const int ld = L.ldS;
  double* R = L.R2;  // scratch factor rows R[t][col] (col = sampled index)
  __shared__ int s_pos2col[64];
  __shared__ double s_sqr[64];
  if (c < 64) {
    const bool act = c < s;
    double d = act ? L.G[c * ld + c] : 0.0;
    int pos = c;
    int rank = s;
    for (int j = 0; j < s; ++j) {
      double bv = -1.0;
      int bp = 0x7fffffff, bc = -1;
      if (act && pos >= j && (d > bv || (d == bv && pos < bp))) {
        bv = d;
        bp = pos;
        bc = c;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int op = __shfl_xor(bp, o, 64);
        const int oc = __shfl_xor(bc, o, 64);
        if (ov > bv || (ov == bv && op < bp)) {
          bv = ov;
          bp = op;
          bc = oc;
        }
      }
      if (!(bv > 0.0)) {
        rank = j;
        break;
      }
      const int P = bc;
      const double rjj = sqrt(bv);
      if (pos == j) pos = bp;
      if (c == P) pos = j;
      if (act) {
        double val = 0.0;
        if (c == P) {
          val = rjj;
        } else if (pos > j) {
          double acc = L.G[P * ld + c];
          for (int t = 0; t < j; ++t) acc = fma(-R[t * ld + P], R[t * ld + c], acc);
          val = acc / rjj;
          d = fma(-val, val, d);
        }
        R[j * ld + c] = val;
      }
    }
    for (int j = rank; j < s; ++j)
      if (act) R[j * ld + c] = 0.0;
    if (act) s_pos2col[pos] = c;
    if (c == 0) L.sc[SC_RANK] = rank;
  }
  __syncthreads();
  if (c < s) {
    double e = 0.0;
    for (int q = 0; q < s; ++q) e = fma(R[c * ld + q], R[c * ld + q], e);
    s_sqr[c] = e;
  }
  __syncthreads();
  if (c == 0) {
    const int rank = L.sc[SC_RANK];
    double total = 0.0;
    for (int t = 0; t < s; ++t) total += s_sqr[s - 1 - t];
    const double thr = total / (double)(filter_mode == 0 ? s : b);
    double csum = 0.0;
    int k = 0;
    for (; k < s; ++k) {
      csum += s_sqr[s - 1 - k];
      if (csum >= thr) break;
    }
    int bt = s - k;
    if (rank == 0 || !(total > 0.0)) {
      bt = 0;
      L.sc[SC_STATUS] = 1;
    }
    if (bt > rank) bt = rank;
    L.sc[SC_BT] = bt;
  }
  __syncthreads();
  const int bt = L.sc[SC_BT];
  if (c < bt) {
    const int col = s_pos2col[c];
    L.Jkpos[c] = col;
    L.Jk[c] = L.J[col];
    for (int r = 0; r < bt; ++r) L.Rb[r * ld + c] = (r <= c) ? R[r * ld + col] : 0.0;
  }
  __syncthreads();
  bq_upper_inverse(L.Rb, L.Rinv, bt, ld);
  // symmetric difference of the kept set and the block positions [i, i + bt)
  if (c < bt) {
    const int v = L.Jk[c];
    int rk = 0;
    for (int q = 0; q < bt; ++q) rk += (L.Jk[q] < v) ? 1 : 0;
    L.found[rk] = v;  // ascending kept indices (reuses the sampler's list)
    const int q = i + c;
    int member = 0;
    for (int a = 0; a < bt; ++a) member |= (L.Jk[a] == q) ? 1 : 0;
    L.isnew[c] = !member;  // position i + c is vacated by a column that leaves
  }
  __syncthreads();
  if (c == 0) {
    int nin = 0, nout = 0;
    for (int a = 0; a < bt; ++a) nin += (L.found[a] >= i + bt) ? 1 : 0;
    for (int a = 0; a < bt; ++a) {
      if (L.isnew[a]) L.Jout[nout++] = i + a;
    }
    for (int a = 0; a < nin; ++a) L.Jin[a] = L.found[bt - nin + a];
    if (nin != nout) L.sc[SC_STATUS] = 3;
    L.sc[SC_NSWAP] = nin;
  }
  __syncthreads();
}

// G2 = R2^T R2 (G2 in L.G2), R2inv, and Rn = R2 * Rb (bt x bt), one wave.
__device__ void bq_chol2(const BqLds& L, int bt) {
  const int c = threadIdx.x;
  const int ld = L.ldS;
  if (c < 64) {
    const bool act = c < bt;
    double d = act ? L.G2[c * ld + c] : 0.0;
    bool bad = false;
    for (int j = 0; j < bt; ++j) {
      const double dj = __shfl(d, j, 64);
      const double rjj = sqrt(dj > 0.0 ? dj : 1.0);
      bad |= !(dj > 0.0);
      if (act) {
        double val = 0.0;
        if (c == j) {
          val = rjj;
        } else if (c > j) {
          double acc = L.G2[j * ld + c];
          for (int t = 0; t < j; ++t) acc = fma(-L.R2[t * ld + j], L.R2[t * ld + c], acc);
          val = acc / rjj;
          d = fma(-val, val, d);
        }
        L.R2[j * ld + c] = val;
      }
    }
    if (c == 0 && bad) L.sc[SC_STATUS] = 2;
    if (act) {
      for (int r = 0; r < bt; ++r) {
        double acc = 0.0;
        for (int l = r; l <= c; ++l) acc = fma(L.R2[r * ld + l], L.Rb[l * ld + c], acc);
        L.Rn[r * ld + c] = (r <= c) ? acc : 0.0;
      }
    }
  }
  __syncthreads();
  bq_upper_inverse(L.R2, L.R2inv, bt, ld);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// The persistent kernel
// ---------------------------------------------------------------------------
template <int RT>
__global__ __launch_bounds__(BQ_NT) __attribute__((amdgpu_waves_per_eu(2))) void k_colid_batched(const RbrpBatchProblem* __restrict__ probs,
                                                         const int* __restrict__ order, int nprob,
                                                         int* queue, RbrpBatchResult* results,
                                                         RbrpBatchShape shape) {
  extern __shared__ double lds_raw[];
  const BqLds L = bq_carve(lds_raw, shape.n_max, shape.b);
  const int tid = threadIdx.x;
  const int b = shape.b, ld = L.ldS;
  char* slot = (char*)shape.ws + (size_t)blockIdx.x * shape.slot_bytes;
  double* Aw = (double*)slot;
  double* R = (double*)(slot + shape.off_R);
  double* Qbuf = (double*)(slot + shape.off_Q);
  double* Cbuf = (double*)(slot + shape.off_C);
  double* Tbuf = (double*)(slot + shape.off_T);
  const int64_t ldr = shape.ldr, ldc = shape.ldr;

  for (;;) {
    if (tid == 0) L.sc[SC_PROB] = atomicAdd(queue, 1);
    __syncthreads();
    const int qi = L.sc[SC_PROB];
    __syncthreads();
    if (qi >= nprob) break;  // uniform: every workgroup leaves once the queue is drained
    const int p = order[qi];
    const RbrpBatchProblem P = probs[p];
    const int m = (int)P.m, n = (int)P.n;
    const int mp = (m + 15) & ~15;
    const int pmn = min(m, n);
    const int kmax = shape.max_rank > 0 ? min(shape.max_rank, pmn) : pmn;
    if (tid == 0) {
      L.sc[SC_STATUS] = 0;
    }
    for (int j = tid; j < n; j += BQ_NT) P.perm[j] = j;
    // ---- copy + squared column norms (wave per column) -------------------
    {
      const int lane = tid & 63, w = tid >> 6;
      for (int j = w; j < n; j += 4) {
        const double* a = P.A + (int64_t)j * P.lda;
        double* aw = Aw + (int64_t)j * mp;
        double acc = 0.0;
        // four rows per lane and iteration, loaded before they are stored
        for (int r0 = lane; r0 < mp; r0 += 256) {
          double v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = r0 + 64 * u;
            v[u] = r < m ? a[r] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = r0 + 64 * u;
            if (r < mp) {
              aw[r] = v[u];
              acc = fma(v[u], v[u], acc);
            }
          }
        }
        acc = rbrp_wave_sum(acc);
        if (lane == 0) L.diags[j] = acc;
      }
    }
    __syncthreads();
    double part0 = 0.0;
    for (int j = tid; j < n; j += BQ_NT) part0 += L.diags[j];
    const double N0 = rbrp_block_sum(part0, L.red);
    const uint64_t base[4] = {P.stream_offset, 0, 0, 0};
    unsigned long long draws = 0;
    double err = (N0 > 0.0) ? 1.0 : 0.0;
    int i = 0, nb = 0;
    bool bad = !(N0 == N0) || N0 > 1.7e308;
    while (!bad && i < kmax && N0 > 0.0) {
      // ---- sample ---------------------------------------------------------
      const int s = bq_sample(L, i, n, kmax, b, P.seed, base, draws);
      if (L.sc[SC_STATUS]) {
        bad = true;
        break;
      }
      if (s <= 0) break;
      // ---- second Gram-Schmidt pass: C = Q^T V, R(0:i, J) += C, V -= Q C ----
      if (i > 0) {
        bq_tn(i, s, mp, [&](int r) { return (const double*)(Aw + (int64_t)r * mp); },
              [&](int c) { return (const double*)(Aw + (int64_t)L.J[c] * mp); }, L.part,
              [&](int r, int c, double v) {
                Cbuf[r + c * ldc] = v;
                R[r + (int64_t)L.J[c] * ldr] += v;
              });
        bq_nn<true, false, 1>(
            mp, s, i, [&](int k) { return (const double*)(Aw + (int64_t)k * mp); },
            [&](int k, int c) { return Cbuf[k + c * ldc]; },
            [&](int c) { return Aw + (int64_t)L.J[c] * mp; }, [&](int, double) {}, L.part);
      }
      // ---- Gram, pivoted Cholesky, filter -------------------------------------
      bq_tn(s, s, mp, [&](int r) { return (const double*)(Aw + (int64_t)L.J[r] * mp); },
            [&](int c) { return (const double*)(Aw + (int64_t)L.J[c] * mp); }, L.part,
            [&](int r, int c, double v) { L.G[r * ld + c] = v; });
      bq_pivchol(L, s, b, shape.filter, i);
      const int bt = L.sc[SC_BT];
      if (L.sc[SC_STATUS] || bt <= 0) {
        bad = L.sc[SC_STATUS] != 0;
        break;
      }
      // ---- CholQR2: Q1 = V(:, Jk) Rb^{-1}, G2 = Q1^T Q1, R2 -------------------
      bq_nn<false, false, 1>(
          mp, bt, bt, [&](int k) { return (const double*)(Aw + (int64_t)L.Jk[k] * mp); },
          [&](int k, int c) { return L.Rinv[k * ld + c]; },
          [&](int c) { return Qbuf + (int64_t)c * mp; }, [&](int, double) {}, L.part);
      bq_tn(bt, bt, mp, [&](int r) { return (const double*)(Qbuf + (int64_t)r * mp); },
            [&](int c) { return (const double*)(Qbuf + (int64_t)c * mp); }, L.part,
            [&](int r, int c, double v) { L.G2[r * ld + c] = v; });
      bq_chol2(L, bt);
      if (L.sc[SC_STATUS]) {
        bad = true;
        break;
      }
      const int nsw = L.sc[SC_NSWAP];
      // ---- symmetric-difference permutation ----------------------------------
      if (tid < bt) {
        L.tk[tid] = P.perm[L.Jk[tid]];
        L.dk[tid] = L.diags[L.Jk[tid]];
      }
      if (tid < nsw) {
        L.to[tid] = P.perm[L.Jout[tid]];
        L.dout[tid] = L.diags[L.Jout[tid]];
      }
      // R rows 0..i of the kept / leaving columns into Tbuf
      bq_copy(
          i * (bt + nsw),
          [&](int idx) {
            const int r = idx % i, q = idx / i;
            return R + r + (int64_t)(q < bt ? L.Jk[q] : L.Jout[q - bt]) * ldr;
          },
          [&](int idx) { return Tbuf + idx % i + (int64_t)(idx / i) * ldr; });
      __syncthreads();
      if (tid < nsw) {
        P.perm[L.Jin[tid]] = L.to[tid];
        L.diags[L.Jin[tid]] = L.dout[tid];
      }
      bq_copy(
          i * nsw, [&](int idx) { return Tbuf + idx % i + (int64_t)(bt + idx / i) * ldr; },
          [&](int idx) { return R + idx % i + (int64_t)L.Jin[idx / i] * ldr; });
      // working-matrix columns leaving the block slots move to the vacated ones
      // (Jin and Jout are disjoint)
      bq_copy(
          mp * nsw, [&](int idx) { return Aw + idx % mp + (int64_t)L.Jout[idx / mp] * mp; },
          [&](int idx) { return Aw + idx % mp + (int64_t)L.Jin[idx / mp] * mp; });
      __syncthreads();
      if (tid < bt) {
        P.perm[i + tid] = L.tk[tid];
        L.diags[i + tid] = L.dk[tid];
      }
      bq_copy(
          i * bt, [&](int idx) { return Tbuf + idx % i + (int64_t)(idx / i) * ldr; },
          [&](int idx) { return R + idx % i + (int64_t)(i + idx / i) * ldr; });
      // new diagonal block of R, zeros left of it in its rows
      for (int idx = tid; idx < bt * bt; idx += BQ_NT) {
        const int r = idx % bt, q = idx / bt;
        R[(i + r) + (int64_t)(i + q) * ldr] = L.Rn[r * ld + q];
      }
      for (int idx = tid; idx < bt * i; idx += BQ_NT) {
        const int r = idx % bt, q = idx / bt;
        R[(i + r) + (int64_t)q * ldr] = 0.0;
      }
      __syncthreads();
      // ---- Q_b = Q1 R2^{-1} into the block's columns ---------------------------
      bq_nn<false, false, 1>(
          mp, bt, bt, [&](int k) { return (const double*)(Qbuf + (int64_t)k * mp); },
          [&](int k, int c) { return L.R2inv[k * ld + c]; },
          [&](int c) { return Aw + (int64_t)(i + c) * mp; }, [&](int, double) {}, L.part);
      // ---- projection W = Q_b^T A_rem into R's block rows, then update -------
      const int nrem = n - i - bt;
      if (bt <= 16 && nrem >= 49) {
        bq_proj_update<RT>(mp, mp, bt, nrem, Aw + (int64_t)i * mp, Aw + (int64_t)(i + bt) * mp,
                       R + i + (int64_t)(i + bt) * ldr, ldr,
                       [&](int c, double v) { L.diags[i + bt + c] = v; });
      } else if (nrem > 0) {
        bq_tn(bt, nrem, mp, [&](int r) { return (const double*)(Aw + (int64_t)(i + r) * mp); },
              [&](int c) { return (const double*)(Aw + (int64_t)(i + bt + c) * mp); }, L.part,
              [&](int r, int c, double v) { R[(i + r) + (int64_t)(i + bt + c) * ldr] = v; });
        bq_nn<true, true, 2>(
            mp, nrem, bt, [&](int k) { return (const double*)(Aw + (int64_t)(i + k) * mp); },
            [&](int k, int c) { return R[(i + k) + (int64_t)(i + bt + c) * ldr]; },
            [&](int c) { return Aw + (int64_t)(i + bt + c) * mp; },
            [&](int c, double v) { L.diags[i + bt + c] = v; }, L.part);
      }
      double pr = 0.0;
      for (int j = i + bt + tid; j < n; j += BQ_NT) pr += L.diags[j];
      const double resid = rbrp_block_sum(pr, L.red);
      err = resid / N0;
      i += bt;
      ++nb;
      if (err <= shape.tol2) break;
    }
    const int k = i;
    // ---- X = R11^{-1} R12 ------------------------------------------------------
    if (!bad && P.X && k > 0 && n > k) {
      const int nc = n - k;
      bq_copy(
          k * nc, [&](int idx) { return R + idx % k + (int64_t)(k + idx / k) * ldr; },
          [&](int idx) { return P.X + idx % k + (int64_t)(idx / k) * P.ldx; });
      __syncthreads();
      for (int rb = ((k - 1) / 16) * 16; rb >= 0; rb -= 16) {
        const int re = min(rb + 16, k);
        if (re < k) {  // X[rb:re, :] -= R[rb:re, re:k] X[re:k, :]
          bq_nn<true, false, 1>(
              re - rb, nc, k - re,
              [&](int kk) { return (const double*)(R + rb + (int64_t)(re + kk) * ldr); },
              [&](int kk, int c) { return P.X[(re + kk) + (int64_t)c * P.ldx]; },
              [&](int c) { return P.X + rb + (int64_t)c * P.ldx; }, [&](int, double) {}, L.part);
        }
        // the diagonal triangle, one column per thread
        for (int c = tid; c < nc; c += BQ_NT) {
          double* xc = P.X + (int64_t)c * P.ldx;
          double x[16];
#pragma unroll
          for (int q = 0; q < 16; ++q) x[q] = (rb + q < re) ? xc[rb + q] : 0.0;
#pragma unroll
          for (int q = 15; q >= 0; --q) {
            if (rb + q < re) {
              double v = x[q];
#pragma unroll
              for (int l = q + 1; l < 16; ++l)
                if (rb + l < re) v = fma(-R[(rb + q) + (int64_t)(rb + l) * ldr], x[l], v);
              x[q] = v / R[(rb + q) + (int64_t)(rb + q) * ldr];
            }
          }
#pragma unroll
          for (int q = 0; q < 16; ++q)
            if (rb + q < re) xc[rb + q] = x[q];
        }
        __syncthreads();
      }
    }
    if (tid == 0) {
      RbrpBatchResult res;
      res.k = k;
      res.nblocks = nb;
      res.err_final = sqrt(err > 0.0 ? err : 0.0);
      res.draws = draws;
      res.status = bad ? (L.sc[SC_STATUS] ? L.sc[SC_STATUS] : 5) : 0;
      res.pad = 0;
      results[p] = res;
    }
    __syncthreads();
  }
}

// The instantiation for a launch shape's update row tiles (1, 2 or 4), or null.
static const void* batched_kernel(int tile) {
  switch (tile) {
    case 1: return (const void*)k_colid_batched<1>;
    case 2: return (const void*)k_colid_batched<2>;
    case 4: return (const void*)k_colid_batched<4>;
    default: return nullptr;
  }
}

__global__ void k_queue_reset(int* queue) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *queue = 0;
}

hipError_t rbrp_launch_colid_batched(hipStream_t st, int grid, const RbrpBatchProblem* probs,
                                     const int* order, int nprob, int* queue,
                                     RbrpBatchResult* results, const RbrpBatchShape& shape) {
  if (grid <= 0 || nprob <= 0) return hipSuccess;
  if (shape.b < 1 || shape.b > RBRP_BATCH_MAXB || shape.n_max > RBRP_BATCH_MAXN)
    return hipErrorInvalidValue;
  const size_t lds = rbrp_batched_lds_bytes(shape.n_max, shape.b);
  if (!batched_kernel(shape.tile)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, queue);
  const dim3 g((unsigned)grid), t(BQ_NT);
  switch (shape.tile) {
    case 1: hipLaunchKernelGGL(k_colid_batched<1>, g, t, lds, st, probs, order, nprob, queue, results, shape); break;
    case 2: hipLaunchKernelGGL(k_colid_batched<2>, g, t, lds, st, probs, order, nprob, queue, results, shape); break;
    default: hipLaunchKernelGGL(k_colid_batched<4>, g, t, lds, st, probs, order, nprob, queue, results, shape); break;
  }
  return hipGetLastError();
}

size_t rbrp_batched_lds_bytes(int n_max, int b) { return bq_lds_bytes(n_max, b); }

int rbrp_batched_max_blocks_per_cu(int n_max, int b, int tile) {
  const void* fn = batched_kernel(tile);
  int nb = 0;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, BQ_NT, bq_lds_bytes(n_max, b)) !=
                 hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

hipError_t rbrp_batched_prepare(int n_max, int b) {
  for (int tile : {1, 2, 4}) {
    const hipError_t e = hipFuncSetAttribute(batched_kernel(tile),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)bq_lds_bytes(n_max, b));
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
