// Blocked column-pivoted Householder QR, the organisation of LAPACK dgeqp3
// with its panel routine dlaqps (the factorization MAGMA's dgeqp3_gpu runs in
// gpu/randCPQR_OS.cu:64 and gpu/id_cpqr.cu:57).
//
// Panels of QP_NB columns.  Column step j of a panel starting at k0 (jb = j -
// k0 columns of the panel done):
//   * pivot: the first maximum of the partial column norms vn1[j:cols); the
//     column swaps with j (all rows), so do the rows j and p of F (its first
//     jb columns), and p takes j's vn1 / vn2;
//   * column update: W(j:, j) -= W(j:, k0:j) F(j, 0:jb)^T (the panel's
//     earlier reflectors, not yet applied to the trailing matrix);
//   * reflector of W(j:, j) with the dlarfg conventions of the unblocked
//     kernels: beta = -sign(alpha) ||x||, tau = (beta - alpha) / beta,
//     v = x(j+1:) / (alpha - beta), beta on the diagonal;
//   * F(c, jb) = tau W(j:, c)^T v + F(c, 0:jb) auxv for c > j, with auxv =
//     -tau W(j:, k0:j)^T v (v(j) = 1);
//   * row update: W(j, c) -= W(j, k0:j) F(c, 0:jb)^T + F(c, jb) for c > j;
//   * norm downdate (while j + 1 < min(rows, cols)): t = |W(j, c)| / vn1[c],
//     t = max(0, (1 + t)(1 - t)); if t (vn1[c] / vn2[c])^2 <= sqrt(eps) the
//     norm is recomputed, vn1[c] = vn2[c] = ||W(j+1:, c)|| of the column
//     with the panel's reflectors applied, else vn1[c] *= sqrt(t).
// After a panel of kb steps: W(r0:, r0:) -= W(r0:, k0:r0) F(r0:, 0:kb)^T with
// r0 = k0 + kb on the MFMA trailing-update kernel.
//
// A column whose norm must be recomputed: dlaqps ends the panel after the
// step and takes the norm from the updated trailing rows after the panel's
// GEMM.  Tall matrices (rows > QP_SMALL_ROWS) do the same: the column steps
// of a panel are queued, a step finding the panel ended does nothing, and
// the host reads the panel width once per panel.  Short matrices instead
// form the column's updated values on the fly (W(i, c) - W(i, k0:j+1)
// F(c, 0:jb+1)^T for i > j, the same quantity), so that every panel is QP_NB
// wide and the whole factorization is queued without a host round trip (a
// short matrix has many cheap steps, where the round trips would dominate).
// Two launches per column step when the column fits one workgroup (rows <=
// QP_SMALL_ROWS), five when the column work is spread over row chunks.
#include <algorithm>
#include <climits>
#include <cmath>

#include "rbrp_kernels.h"

#define QP_NB 32
#define QP_SMALL_ROWS 4096
#define QP_CHUNK 512   // rows per workgroup of the tall column kernels
#define QP_COLS 8      // trailing columns per workgroup of the update kernel
#define QP_RCH 4096    // rows per chunk of the tall dots kernel
// sqrt of LAPACK's relative machine precision dlamch('E') = 2^-53
#define QP_TOL3Z 1.0536712127723509e-08

struct QpState {
  int stop;      // the panel has ended: its remaining column steps do nothing
  int marked;    // the last column step marked a column for recomputation
  int kb;        // column steps run in the panel
  int p;         // the current step's pivot (tall path)
  double alpha;  // x(j) after the column update (tall path)
};

namespace {

// Whether column step jb runs (thread 0 decides, every thread reads *go): not
// after the panel has ended, which a marked column does.
__device__ inline void qp_gate(QpState* st, int jb, int* go) {
  if (threadIdx.x == 0) {
    int g = !st->stop;
    if (g && st->marked) {
      st->stop = 1;
      g = 0;
    }
    if (g) st->kb = jb + 1;
    *go = g;
  }
  __syncthreads();
}

// First maximum of vn1[j:cols) over a 1024-thread workgroup: the pivot,
// then the swaps of the columns (all rows, unless `rows_too` is false), the F
// rows, jpvt and the copies of the norms.
__device__ inline int64_t qp_pivot(int64_t rows, int64_t cols, int64_t j, int jb, double* W,
                                   int64_t ld, double* vn1, double* vn2, int* jpvt, double* F,
                                   int64_t ldf, bool rows_too, double* rv, int* ri, int* s_p) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double bv = -1.0;
  int bi = INT_MAX;
  for (int64_t c = j + tid; c < cols; c += 1024) {
    const double d = vn1[c];
    if (d > bv) {  // increasing c per thread: keeps the first maximum
      bv = d;
      bi = (int)c;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(bv, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ov > bv || (ov == bv && oi < bi)) {
      bv = ov;
      bi = oi;
    }
  }
  if (lane == 0) {
    rv[wid] = bv;
    ri[wid] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    double b = rv[0];
    int i = ri[0];
    for (int w = 1; w < 16; ++w)
      if (rv[w] > b || (rv[w] == b && ri[w] < i)) {
        b = rv[w];
        i = ri[w];
      }
    *s_p = (i == INT_MAX) ? (int)j : i;
  }
  __syncthreads();
  const int64_t p = *s_p;
  if (p != j) {
    if (rows_too)
      for (int64_t i = tid; i < rows; i += 1024) {
        const double a = W[i + j * ld];
        W[i + j * ld] = W[i + p * ld];
        W[i + p * ld] = a;
      }
    for (int q = tid; q < jb; q += 1024) {
      const double a = F[j + q * ldf];
      F[j + q * ldf] = F[p + q * ldf];
      F[p + q * ldf] = a;
    }
    if (tid == 0) {
      const int t = jpvt[j];
      jpvt[j] = jpvt[p];
      jpvt[p] = t;
      vn1[p] = vn1[j];
      vn2[p] = vn2[j];
    }
  }
  __syncthreads();
  return p;
}

// dlarfg on alpha and the squared norm xn2 of x(j+1:).
__device__ inline void qp_larfg(double alpha, double xn2, double* t, double* beta, double* scal) {
  *t = 0.0;
  *beta = alpha;
  *scal = 0.0;
  if (xn2 != 0.0) {
    const double nrm = sqrt(alpha * alpha + xn2);
    *beta = (alpha >= 0.0) ? -nrm : nrm;
    *t = (*beta - alpha) / *beta;
    *scal = 1.0 / (alpha - *beta);
  }
}

}  // namespace

// Column step, one 1024-thread workgroup (rows <= QP_SMALL_ROWS): pivot,
// column update, reflector, auxv.
__global__ __launch_bounds__(1024) void k_qp_col_small(int64_t rows, int64_t cols, int64_t j,
                                                       int64_t k0, int jb, double* __restrict__ W,
                                                       int64_t ld, double* __restrict__ vn1,
                                                       double* __restrict__ vn2, int* __restrict__ jpvt,
                                                       double* __restrict__ F, int64_t ldf,
                                                       double* __restrict__ tau,
                                                       double* __restrict__ auxv) {
  __shared__ double rv[16], red[16], fj[QP_NB];
  __shared__ int ri[16], s_p;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  (void)qp_pivot(rows, cols, j, jb, W, ld, vn1, vn2, jpvt, F, ldf, true, rv, ri, &s_p);
  if (tid < jb) fj[tid] = F[j + tid * ldf];
  __syncthreads();
  double* x = W + j * ld;
  double acc = 0.0;
  for (int64_t i = j + tid; i < rows; i += 1024) {
    double v = x[i];
    for (int q = 0; q < jb; ++q) v -= W[i + (k0 + q) * ld] * fj[q];
    x[i] = v;
    if (i > j) acc = fma(v, v, acc);
  }
  const double xn2 = rbrp_block_sum(acc, red);
  const double alpha = x[j];
  __syncthreads();
  double t, beta, scal;
  qp_larfg(alpha, xn2, &t, &beta, &scal);
  if (xn2 != 0.0)
    for (int64_t i = j + 1 + tid; i < rows; i += 1024) x[i] *= scal;
  __syncthreads();
  if (tid == 0) {
    tau[j] = t;
    x[j] = beta;
  }
  // auxv[q] = -tau (W(j, k0+q) + sum_{i>j} W(i, k0+q) v(i)), a wave per q
  for (int q = wid; q < jb; q += 16) {
    const double* wc = W + (k0 + q) * ld;
    double a = 0.0;
    for (int64_t i = j + 1 + lane; i < rows; i += 64) a = fma(wc[i], x[i], a);
    a = rbrp_wave_sum(a);
    if (lane == 0) auxv[q] = -t * (wc[j] + a);
  }
}

// Tall column step, part 1 (one workgroup): the pivot and the swaps of F,
// jpvt and the norms (the columns are swapped by part 2).
__global__ __launch_bounds__(1024) void k_qp_pivot(int64_t rows, int64_t cols, int64_t j, int jb,
                                                   double* __restrict__ W, int64_t ld,
                                                   double* __restrict__ vn1, double* __restrict__ vn2,
                                                   int* __restrict__ jpvt, double* __restrict__ F,
                                                   int64_t ldf, QpState* st) {
  __shared__ double rv[16];
  __shared__ int ri[16], s_p, s_go;
  qp_gate(st, jb, &s_go);
  if (!s_go) return;
  const int64_t p = qp_pivot(rows, cols, j, jb, W, ld, vn1, vn2, jpvt, F, ldf, false, rv, ri, &s_p);
  if (threadIdx.x == 0) st->p = (int)p;
}

// Tall column step, part 2 (QP_CHUNK rows per workgroup): the column swap,
// the column update, and per workgroup the partial sum of x(j+1:)^2 and the
// partial dots sum_{i>j} W(i, k0+q) x(i) to part[wg (QP_NB + 1) + (0, 1 + q)].
__global__ __launch_bounds__(256) void k_qp_col_tall(int64_t rows, int64_t j, int64_t k0, int jb,
                                                     double* __restrict__ W, int64_t ld,
                                                     const double* __restrict__ F, int64_t ldf,
                                                     double* __restrict__ part, QpState* st) {
  __shared__ double fj[QP_NB], red[4];
  if (st->stop) return;  // uniform
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t p = st->p;
  const int64_t r0 = (int64_t)blockIdx.x * QP_CHUNK, r1 = min(rows, r0 + QP_CHUNK);
  double* x = W + j * ld;
  if (p != j)
    for (int64_t i = r0 + tid; i < r1; i += 256) {
      const double a = x[i];
      x[i] = W[i + p * ld];
      W[i + p * ld] = a;
    }
  if (tid < jb) fj[tid] = F[j + tid * ldf];
  __syncthreads();
  double acc = 0.0;
  for (int64_t i = max(r0, j) + tid; i < r1; i += 256) {
    double v = x[i];
    for (int q = 0; q < jb; ++q) v -= W[i + (k0 + q) * ld] * fj[q];
    x[i] = v;
    if (i == j) st->alpha = v;
    if (i > j) acc = fma(v, v, acc);
  }
  const double s = rbrp_block_sum(acc, red);  // (its barriers make x visible)
  double* pw = part + (int64_t)blockIdx.x * (QP_NB + 1);
  if (tid == 0) pw[0] = s;
  for (int q = wid; q < jb; q += 4) {
    const double* wc = W + (k0 + q) * ld;
    double a = 0.0;
    for (int64_t i = max(r0, j + 1) + lane; i < r1; i += 64) a = fma(wc[i], x[i], a);
    a = rbrp_wave_sum(a);
    if (lane == 0) pw[1 + q] = a;
  }
}

// Tall column step, part 3 (the same chunks): every workgroup adds the
// partials in one fixed order (per-thread strides, then the workgroup tree),
// forms the reflector and scales its rows of x; workgroup 0 stores tau, beta
// and auxv.
__global__ __launch_bounds__(256) void k_qp_refl_tall(int64_t rows, int64_t j, int64_t k0, int jb,
                                                      double* __restrict__ W, int64_t ld,
                                                      const double* __restrict__ part, int nparts,
                                                      double* __restrict__ tau,
                                                      double* __restrict__ auxv, QpState* st) {
  __shared__ double red[4];
  if (st->stop) return;  // uniform
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double a = 0.0;
  for (int w = tid; w < nparts; w += 256) a += part[(int64_t)w * (QP_NB + 1)];
  const double xn2 = rbrp_block_sum(a, red);
  const double alpha = st->alpha;
  double t, beta, scal;
  qp_larfg(alpha, xn2, &t, &beta, &scal);
  double* x = W + j * ld;
  const int64_t r0 = (int64_t)blockIdx.x * QP_CHUNK, r1 = min(rows, r0 + QP_CHUNK);
  if (xn2 != 0.0)
    for (int64_t i = max(r0, j + 1) + tid; i < r1; i += 256) x[i] *= scal;
  if (blockIdx.x == 0) {
    if (tid == 0) {
      tau[j] = t;
      x[j] = beta;
    }
    for (int q = wid; q < jb; q += 4) {
      double d = 0.0;
      for (int w = lane; w < nparts; w += 64) d += part[(int64_t)w * (QP_NB + 1) + 1 + q];
      d = rbrp_wave_sum(d);
      if (lane == 0) auxv[q] = -t * (W[j + (k0 + q) * ld] + scal * d);
    }
  }
}

// Partial dots sum_{i in [i0, i1)} W(i, c0 + u) v(i) of a 256-thread
// workgroup, u < QP_COLS, reduced into red[u][wave] (columns past cols are
// clamped loads whose sums are never used).
__device__ inline void qp_dots(const double* __restrict__ W, int64_t ld, int64_t cols, int64_t c0,
                               const double* __restrict__ v, int64_t i0, int64_t i1,
                               double (*red)[4]) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double* wc[QP_COLS];
#pragma unroll
  for (int u = 0; u < QP_COLS; ++u) wc[u] = W + min(c0 + u, cols - 1) * ld;  // (clamped: loads issue together)
  double acc[QP_COLS];
#pragma unroll
  for (int u = 0; u < QP_COLS; ++u) acc[u] = 0.0;
  int64_t i = i0 + tid;
  for (; i + 768 < i1; i += 1024) {
    double vv[4], ww[4][QP_COLS];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      vv[h] = v[i + 256 * h];
#pragma unroll
      for (int u = 0; u < QP_COLS; ++u) ww[h][u] = wc[u][i + 256 * h];
    }
#pragma unroll
    for (int h = 0; h < 4; ++h)
#pragma unroll
      for (int u = 0; u < QP_COLS; ++u) acc[u] = fma(ww[h][u], vv[h], acc[u]);
  }
  for (; i < i1; i += 256) {
    const double vi = v[i];
#pragma unroll
    for (int u = 0; u < QP_COLS; ++u) acc[u] = fma(wc[u][i], vi, acc[u]);
  }
#pragma unroll
  for (int u = 0; u < QP_COLS; ++u) {
    const double a = rbrp_wave_sum(acc[u]);
    if (lane == 0) red[u][wid] = a;
  }
}

// Tall matrices: the dots of the F column over row chunks of QP_RCH rows,
// chunk first + y for grid row y (the chunks holding rows below j; grid:
// column groups x chunks), pdot[y * cols + c].
__global__ __launch_bounds__(256) void k_qp_dots(int64_t rows, int64_t cols, int64_t j, int64_t first,
                                                 const double* __restrict__ W, int64_t ld,
                                                 double* __restrict__ pdot, const QpState* st) {
  __shared__ double red[QP_COLS][4];
  if (st->stop) return;  // uniform
  const int64_t c0 = j + 1 + (int64_t)blockIdx.x * QP_COLS;
  const int64_t ch = first + blockIdx.y;
  const int64_t i0 = max(j + 1, ch * QP_RCH), i1 = min(rows, (ch + 1) * QP_RCH);
  qp_dots(W, ld, cols, c0, W + j * ld, i0, i1, red);
  __syncthreads();
  const int tid = threadIdx.x;
  if (tid < QP_COLS && c0 + tid < cols)
    pdot[(int64_t)blockIdx.y * cols + c0 + tid] = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
}

// F column, row update and norm downdate for trailing columns c > j,
// QP_COLS columns per 256-thread workgroup.  A column whose downdate is
// rejected gets its norm recomputed here from the updated values (redo_here)
// or is marked, ending the panel.  The dots W(j+1:, c)^T v are formed here
// (nrc = 0) or summed from k_qp_dots' nrc chunk partials.
__global__ __launch_bounds__(256) void k_qp_update(int64_t rows, int64_t cols, int64_t j, int64_t k0,
                                                   int jb, double* __restrict__ W, int64_t ld,
                                                   double* __restrict__ vn1, double* __restrict__ vn2,
                                                   double* __restrict__ F, int64_t ldf,
                                                   const double* __restrict__ tau,
                                                   const double* __restrict__ auxv,
                                                   const double* __restrict__ pdot, int nrc,
                                                   int downdate, int redo_here, int* __restrict__ mark,
                                                   QpState* st) {
  static_assert(QP_COLS * QP_NB == 256, "one F element per thread");
  __shared__ double red[QP_COLS][4], s_av[QP_NB], s_wj[QP_NB], s_f[QP_COLS][QP_NB];
  __shared__ int s_redo[QP_COLS];
  if (st->stop) return;  // uniform
  const int tid = threadIdx.x;
  const int64_t c0 = j + 1 + (int64_t)blockIdx.x * QP_COLS;
  {
    // F(c0 + u, q), q < jb, staged in LDS (loaded while the dots run)
    const int u = tid / QP_NB, q = tid % QP_NB;
    s_f[u][q] = (q < jb) ? F[min(c0 + u, cols - 1) + q * ldf] : 0.0;
  }
  if (tid < jb) {
    s_av[tid] = auxv[tid];
    s_wj[tid] = W[j + (k0 + tid) * ld];
  }
  const double* v = W + j * ld;
  if (nrc == 0) qp_dots(W, ld, cols, c0, v, j + 1, rows, red);
  __syncthreads();
  if (tid < QP_COLS) {
    const int64_t c = c0 + tid;
    int redo = 0;
    if (c < cols) {
      const double t = tau[j];
      const double wjc = W[j + c * ld];
      double d = 0.0;
      if (nrc == 0) {
        d = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
      } else {
        for (int rc = 0; rc < nrc; ++rc) d += pdot[(int64_t)rc * cols + c];
      }
      const double dot = wjc + d;
      double f = t * dot;
      for (int q = 0; q < jb; ++q) f = fma(s_f[tid][q], s_av[q], f);
      F[c + jb * ldf] = f;
      double r = wjc;
      for (int q = 0; q < jb; ++q) r -= s_wj[q] * s_f[tid][q];
      s_f[tid][jb] = f;
      r -= f;
      W[j + c * ld] = r;
      if (downdate) {
        const double n1 = vn1[c];
        if (n1 != 0.0) {
          double tt = fabs(r) / n1;
          tt = fmax(0.0, (1.0 + tt) * (1.0 - tt));
          const double ratio = n1 / vn2[c];
          if (tt * ratio * ratio > QP_TOL3Z) {
            vn1[c] = n1 * sqrt(tt);
          } else if (redo_here) {
            redo = 1;
          } else {
            mark[c] = 1;
            st->marked = 1;
          }
        }
      }
    }
    s_redo[tid] = redo;
  }
  __syncthreads();
  // recomputation: ||W(j+1:, c) - W(j+1:, k0:j+1) F(c, 0:jb+1)^T|| (the
  // reflector of column j is W(j+1:, j) below its unit diagonal)
  for (int u = 0; u < QP_COLS; ++u) {
    if (!s_redo[u]) continue;  // uniform
    const int64_t c = c0 + u;
    double sq = 0.0;
    for (int64_t r = j + 1 + tid; r < rows; r += 256) {
      double val = W[r + c * ld];
      for (int q = 0; q <= jb; ++q) val -= W[r + (k0 + q) * ld] * s_f[u][q];
      sq = fma(val, val, sq);
    }
    const double s = rbrp_block_sum(sq, &red[0][0]);
    if (tid == 0) {
      vn1[c] = sqrt(s);
      vn2[c] = vn1[c];
    }
    __syncthreads();
  }
}

// After a panel of the tall path: the marked columns c >= r0 get vn1 = vn2 =
// ||W(r0:, c)||; workgroup 0 also clears the panel state.
__global__ __launch_bounds__(256) void k_qp_recompute(int64_t rows, int64_t cols, int64_t r0,
                                                      const double* __restrict__ W, int64_t ld,
                                                      double* __restrict__ vn1,
                                                      double* __restrict__ vn2, int* __restrict__ mark,
                                                      QpState* st) {
  __shared__ double red[4];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->stop = 0;
    st->marked = 0;
    st->kb = 0;
  }
  const int64_t c = r0 + blockIdx.x;
  if (c >= cols || !mark[c]) return;  // uniform
  double acc = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < rows; i += 256) acc = fma(W[i + c * ld], W[i + c * ld], acc);
  const double s = rbrp_block_sum(acc, red);
  if (threadIdx.x == 0) {
    vn1[c] = sqrt(s);
    vn2[c] = vn1[c];
    mark[c] = 0;
  }
}

// vn1 = vn2 = sqrt of the squared norms in vn1, jpvt = identity, no marks, a
// clear panel state.
__global__ void k_qp_init(int64_t cols, double* __restrict__ vn1, double* __restrict__ vn2,
                          int* __restrict__ jpvt, int* __restrict__ mark, QpState* st) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) {
    st->stop = 0;
    st->marked = 0;
    st->kb = 0;
  }
  if (c < cols) {
    vn1[c] = sqrt(vn1[c]);
    vn2[c] = vn1[c];
    jpvt[c] = (int)c;
    mark[c] = 0;
  }
}

namespace {

struct QpWork {
  double *F, *vn1, *vn2, *auxv, *part, *pdot;
  int* mark;
  QpState* st;
};

inline size_t qp_align(size_t x) { return (x + 255) / 256 * 256; }

int64_t qp_nparts(int64_t rows) { return (rows + QP_CHUNK - 1) / QP_CHUNK; }
// row chunks of the dots kernel (0: the update kernel forms the dots itself)
int64_t qp_nrc(int64_t rows) { return rows <= QP_SMALL_ROWS ? 0 : (rows + QP_RCH - 1) / QP_RCH; }

QpWork qp_carve(void* work, int64_t rows, int64_t cols) {
  char* p = (char*)work;
  QpWork w;
  w.F = (double*)p;
  p += qp_align(sizeof(double) * cols * QP_NB);
  w.vn1 = (double*)p;
  p += qp_align(sizeof(double) * cols);
  w.vn2 = (double*)p;
  p += qp_align(sizeof(double) * cols);
  w.auxv = (double*)p;
  p += qp_align(sizeof(double) * QP_NB);
  w.part = (double*)p;
  p += qp_align(sizeof(double) * qp_nparts(rows) * (QP_NB + 1));
  w.pdot = (double*)p;
  p += qp_align(sizeof(double) * qp_nrc(rows) * cols);
  w.mark = (int*)p;
  p += qp_align(sizeof(int) * cols);
  w.st = (QpState*)p;
  return w;
}

}  // namespace

size_t rbrp_qrcp_work_bytes(int64_t rows, int64_t cols) {
  return qp_align(sizeof(double) * cols * QP_NB) + 2 * qp_align(sizeof(double) * cols) +
         qp_align(sizeof(double) * QP_NB) + qp_align(sizeof(double) * qp_nparts(rows) * (QP_NB + 1)) +
         qp_align(sizeof(double) * qp_nrc(rows) * cols) + qp_align(sizeof(int) * cols) +
         qp_align(sizeof(QpState));
}

hipError_t rbrp_qrcp(hipStream_t st, int64_t rows, int64_t cols, double* W, int64_t ld, int* jpvt,
                     double* tau, void* work, int* host_kb, int* panels) {
  if (panels) *panels = 0;
  if (rows <= 0 || cols <= 0) return hipSuccess;
  if (ld < rows || !W || !jpvt || !tau || !work || !host_kb) return hipErrorInvalidValue;
  const QpWork w = qp_carve(work, rows, cols);
  const int64_t ldf = cols, q = std::min(rows, cols), lastrk = std::min(rows, cols);
  const bool small = rows <= QP_SMALL_ROWS;
  const int64_t nparts = qp_nparts(rows), nrc_all = qp_nrc(rows);
  hipError_t e;
  if ((e = rbrp_launch_colnorm2(st, rows, cols, W, ld, w.vn1)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_qp_init, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, st, cols, w.vn1,
                     w.vn2, jpvt, w.mark, w.st);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int np = 0;
  int64_t k0 = 0;
  while (k0 < q) {
    const int nbp = (int)std::min<int64_t>(QP_NB, q - k0);
    for (int jb = 0; jb < nbp; ++jb) {
      const int64_t j = k0 + jb;
      if (small) {
        hipLaunchKernelGGL(k_qp_col_small, dim3(1), dim3(1024), 0, st, rows, cols, j, k0, jb, W, ld,
                           w.vn1, w.vn2, jpvt, w.F, ldf, tau, w.auxv);
      } else {
        hipLaunchKernelGGL(k_qp_pivot, dim3(1), dim3(1024), 0, st, rows, cols, j, jb, W, ld, w.vn1,
                           w.vn2, jpvt, w.F, ldf, w.st);
        hipLaunchKernelGGL(k_qp_col_tall, dim3((unsigned)nparts), dim3(256), 0, st, rows, j, k0, jb,
                           W, ld, (const double*)w.F, ldf, w.part, w.st);
        hipLaunchKernelGGL(k_qp_refl_tall, dim3((unsigned)nparts), dim3(256), 0, st, rows, j, k0, jb,
                           W, ld, (const double*)w.part, (int)nparts, tau, w.auxv, w.st);
      }
      if (j + 1 < cols) {
        const int64_t ntr = cols - j - 1, ncg = (ntr + QP_COLS - 1) / QP_COLS;
        // the row chunks holding rows below j (none past the last row: then
        // the one chunk's sums are empty)
        const int64_t first = std::min((j + 1) / QP_RCH, nrc_all - 1);
        const int64_t nrc = nrc_all ? nrc_all - first : 0;
        if (nrc)
          hipLaunchKernelGGL(k_qp_dots, dim3((unsigned)ncg, (unsigned)nrc), dim3(256), 0, st, rows, cols,
                             j, first, (const double*)W, ld, w.pdot, (const QpState*)w.st);
        hipLaunchKernelGGL(k_qp_update, dim3((unsigned)ncg), dim3(256), 0, st, rows, cols, j, k0, jb, W,
                           ld, w.vn1, w.vn2, w.F, ldf, (const double*)tau, (const double*)w.auxv,
                           (const double*)w.pdot, (int)nrc, (j + 1 < lastrk) ? 1 : 0, small ? 1 : 0,
                           w.mark, w.st);
      }
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    // the panel's width: every step (short matrices), or read back (tall)
    int kb = nbp;
    if (!small) {
      if ((e = hipMemcpyAsync(host_kb, &w.st->kb, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
      kb = *host_kb;
      if (kb < 1 || kb > nbp) return hipErrorUnknown;
    }
    const int64_t r0 = k0 + kb;
    if (r0 < rows && r0 < cols) {
      e = rbrp_launch_gemm_trail(st, rows - r0, dim_static(cols - r0), cols - r0, dim_static(kb), kb,
                                 W + r0, ld, k0, nullptr, w.F + r0, ldf, W + r0, ld, r0, nullptr,
                                 nullptr, 0);
      if (e != hipSuccess) return e;
    }
    if (!small) {
      hipLaunchKernelGGL(k_qp_recompute, dim3((unsigned)std::max<int64_t>(1, cols - r0)), dim3(256), 0,
                         st, rows, cols, r0, (const double*)W, ld, w.vn1, w.vn2, w.mark, w.st);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    k0 = r0;
    ++np;
  }
  if (panels) *panels = np;
  return hipSuccess;
}
