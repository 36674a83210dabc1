// Kernels of the sketch-based ID comparators (RandCPQR_OS, RandLUPP_OS and the
// oversampled interpolation matrix): Gaussian sketch generation, Householder
// QR steps with optional column pivoting, LU steps with partial pivoting, and
// column gathers / transposes.  The GEMM Y = G A runs on k_gemm_nn.
#include "rbrp_kernels.h"

#define SK_NT 256

// ---------------------------------------------------------------------------
// Gaussian sketch: out(row, col) = scale * z_t, t = row + col * rows, where
// (z_{2q}, z_{2q+1}) is the Box-Muller pair of uniforms (u_{2q}, u_{2q+1}) of
// the Philox stream (key (seed, 1), counter `offset`):
//   z_{2q} = sqrt(-2 log(1 - u_{2q})) cos(2 pi u_{2q+1}),
//   z_{2q+1} = sqrt(-2 log(1 - u_{2q})) sin(2 pi u_{2q+1}).
// Replaces Random::Gaussian (gpu/random.cu:44-48, Thrust normal_distribution
// on a global minstd stream) with an explicit, reproducible stream.
// ---------------------------------------------------------------------------
__global__ void k_gaussian(uint64_t key0, uint64_t c0, int64_t rows, int64_t cols, double scale,
                           double* __restrict__ out, int64_t ldo) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= rows * cols) return;
  const uint64_t base[4] = {c0, 0, 0, 0};
  const uint64_t q = (uint64_t)t >> 1;
  const double u1 = rbrp_u64_to_unit(rbrp_philox_word(base, key0, 1, 2 * q));
  const double u2 = rbrp_u64_to_unit(rbrp_philox_word(base, key0, 1, 2 * q + 1));
  const double rad = sqrt(-2.0 * log(1.0 - u1));
  double sn, cs;
  sincos(2.0 * 3.14159265358979323846 * u2, &sn, &cs);
  const double z = (t & 1) ? rad * sn : rad * cs;
  const int64_t row = t % rows, col = t / rows;
  out[row + col * ldo] = scale * z;
}

hipError_t rbrp_launch_gaussian(hipStream_t st, uint64_t seed, uint64_t offset, int64_t rows,
                                int64_t cols, double scale, double* out, int64_t ldo) {
  const int64_t total = rows * cols;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gaussian, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, seed, offset,
                     rows, cols, scale, out, ldo);
  return hipGetLastError();
}

// dst(:, c) = src(:, perm[c]) for c < cols, rows rows.
__global__ void k_gather_cols(int64_t rows, const double* __restrict__ src, int64_t lds,
                              const int* __restrict__ perm, double* __restrict__ dst, int64_t ldd) {
  const int64_t r = (int64_t)blockIdx.y * 256 + threadIdx.x;
  const int64_t c = blockIdx.x;
  if (r >= rows) return;
  dst[r + c * ldd] = src[r + (int64_t)perm[c] * lds];
}

hipError_t rbrp_launch_gather_cols(hipStream_t st, int64_t rows, int64_t cols, const double* src,
                                   int64_t lds, const int* perm, double* dst, int64_t ldd) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)cols, (unsigned)((rows + 255) / 256)), dim3(256), 0,
                     st, rows, src, lds, perm, dst, ldd);
  return hipGetLastError();
}

// out(r, j) = Y(j, P[p0 + r]) for r < nr, j < nj: the transposed sketch
// columns of the given positions of the permutation P (RandAdapLUPP).
__global__ void k_gather_cols_t(int64_t nr, const double* __restrict__ Y, int64_t ldy,
                                const int* __restrict__ P, int64_t p0, double* __restrict__ out,
                                int64_t ldo) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (r >= nr) return;
  out[r + j * ldo] = Y[j + (int64_t)P[p0 + r] * ldy];
}

hipError_t rbrp_launch_gather_cols_t(hipStream_t st, int64_t nr, int64_t nj, const double* Y,
                                     int64_t ldy, const int* P, int64_t p0, double* out, int64_t ldo) {
  if (nr <= 0 || nj <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_cols_t, dim3((unsigned)((nr + 255) / 256), (unsigned)nj), dim3(256), 0,
                     st, nr, Y, ldy, P, p0, out, ldo);
  return hipGetLastError();
}

// The row interchanges base + j <-> base + ipiv[j], j = 0 .. nb-1 in order,
// applied to columns 0 .. cols-1 of M (ld ldm) and to the int array P: one
// thread per column (thread `cols` takes P).
__global__ void k_swap_rows(int nb, int64_t base, const int* __restrict__ ipiv, double* M, int64_t ldm,
                            int64_t cols, int* P) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c < cols) {
    double* col = M + c * ldm + base;
    for (int j = 0; j < nb; ++j) {
      const int q = ipiv[j];
      if (q != j) {
        const double t = col[j];
        col[j] = col[q];
        col[q] = t;
      }
    }
  } else if (c == cols && P) {
    int* pp = P + base;
    for (int j = 0; j < nb; ++j) {
      const int q = ipiv[j];
      if (q != j) {
        const int t = pp[j];
        pp[j] = pp[q];
        pp[q] = t;
      }
    }
  }
}

hipError_t rbrp_launch_swap_rows(hipStream_t st, int nb, int64_t base, const int* ipiv, double* M,
                                 int64_t ldm, int64_t cols, int* P) {
  if (nb <= 0) return hipSuccess;
  const int64_t threads = cols + 1;
  hipLaunchKernelGGL(k_swap_rows, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, nb, base,
                     ipiv, M, ldm, cols, P);
  return hipGetLastError();
}

// dst (cols x rows, ldd) = src(0:rows, 0:cols)^T, 32x32 tiles through LDS.
__global__ void k_transpose(int64_t rows, int64_t cols, const double* __restrict__ src, int64_t lds,
                            double* __restrict__ dst, int64_t ldd) {
  __shared__ double tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int q = ty; q < 32; q += 8) {
    const int64_t r = r0 + tx, c = c0 + q;
    tile[q][tx] = (r < rows && c < cols) ? src[r + c * lds] : 0.0;
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int64_t c = c0 + tx, r = r0 + q;  // dst(c, r)
    if (r < rows && c < cols) dst[c + r * ldd] = tile[tx][q];
  }
}

hipError_t rbrp_launch_transpose(hipStream_t st, int64_t rows, int64_t cols, const double* src,
                                 int64_t lds, double* dst, int64_t ldd) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32)),
                     dim3(256), 0, st, rows, cols, src, lds, dst, ldd);
  return hipGetLastError();
}

// R (k x k, ld) <- its strictly upper part with a unit diagonal and zeros
// below: the unit upper-triangular factor L^T of an LU whose transposed block
// shares storage with U^T.
__global__ void k_unit_upper(int64_t k, double* __restrict__ R, int64_t ld) {
  const int64_t c = blockIdx.x;
  for (int64_t r = c + threadIdx.x; r < k; r += blockDim.x) R[r + c * ld] = (r == c) ? 1.0 : 0.0;
}

hipError_t rbrp_launch_unit_upper(hipStream_t st, int64_t k, double* R, int64_t ld) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_unit_upper, dim3((unsigned)k), dim3(256), 0, st, k, R, ld);
  return hipGetLastError();
}

// V (rows_pad x jb, ldv) <- the Householder vectors of a factored panel
// (rows x jb at src, lds): unit diagonal, the stored vectors below it, zeros
// above it and in the pad rows [rows, rows_pad).
__global__ void k_panel_vectors(int64_t rows, int64_t rows_pad, int jb, const double* __restrict__ src,
                                int64_t lds, double* __restrict__ V, int64_t ldv) {
  const int c = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows_pad || c >= jb) return;
  double x = 0.0;
  if (i < rows) x = (i == c) ? 1.0 : (i > c ? src[i + (int64_t)c * lds] : 0.0);
  V[i + (int64_t)c * ldv] = x;
}

hipError_t rbrp_launch_panel_vectors(hipStream_t st, int64_t rows, int64_t rows_pad, int jb,
                                     const double* src, int64_t lds, double* V, int64_t ldv) {
  if (rows_pad <= 0 || jb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_panel_vectors, dim3((unsigned)((rows_pad + 255) / 256), (unsigned)jb), dim3(256), 0,
                     st, rows, rows_pad, jb, src, lds, V, ldv);
  return hipGetLastError();
}

// Tt (jb x jb, ldt) <- T^T for the block reflector H_0 H_1 ... H_{jb-1} =
// I - V T V^T (forward, column-wise: T(i, i) = tau_i, T(0:i, i) = -tau_i
// T(0:i, 0:i) S(0:i, i) with S = V^T V given, ld lds).  One workgroup of 64
// threads, jb <= 64; thread r owns row r of T.
__global__ __launch_bounds__(64) void k_wy_factor(int jb, const double* __restrict__ S, int64_t lds,
                                                  const double* __restrict__ tau,
                                                  double* __restrict__ Tt, int64_t ldt) {
  __shared__ double T[64 * 64];   // T(r, c) at r + 64 c
  __shared__ double Ss[64 * 64];  // S(c, i) at c + 64 i
  const int r = threadIdx.x;
  if (r < jb)
    for (int i = 0; i < jb; ++i) Ss[r + 64 * i] = S[r + (int64_t)i * lds];
  __syncthreads();
  for (int i = 0; i < jb; ++i) {
    const double ti = tau[i];
    if (r < i) {
      double acc = 0.0;
      for (int c = r; c < i; ++c) acc = fma(T[r + 64 * c], Ss[c + 64 * i], acc);
      T[r + 64 * i] = -ti * acc;
    } else if (r == i) {
      T[r + 64 * i] = ti;
    }
    __syncthreads();
  }
  if (r < jb)
    for (int c = 0; c < jb; ++c) Tt[c + (int64_t)r * ldt] = (r <= c) ? T[r + 64 * c] : 0.0;
}

hipError_t rbrp_launch_wy_factor(hipStream_t st, int jb, const double* S, int64_t lds, const double* tau,
                                 double* Tt, int64_t ldt) {
  if (jb <= 0 || jb > 64) return jb <= 0 ? hipSuccess : hipErrorInvalidValue;
  hipLaunchKernelGGL(k_wy_factor, dim3(1), dim3(64), 0, st, jb, S, lds, tau, Tt, ldt);
  return hipGetLastError();
}

// vn[c] = sum_{i < rows} W(i, c)^2, one workgroup per column.
__global__ __launch_bounds__(SK_NT) void k_colnorm2(int64_t rows, const double* __restrict__ W,
                                                    int64_t ld, double* __restrict__ vn) {
  __shared__ double red[SK_NT / 64];
  const int64_t c = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < rows; i += SK_NT) acc = fma(W[i + c * ld], W[i + c * ld], acc);
  const double t = rbrp_block_sum(acc, red);
  if (threadIdx.x == 0) vn[c] = t;
}

hipError_t rbrp_launch_colnorm2(hipStream_t st, int64_t rows, int64_t cols, const double* W,
                                int64_t ld, double* vn) {
  if (cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_colnorm2, dim3((unsigned)cols), dim3(SK_NT), 0, st, rows, W, ld, vn);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Unpivoted Householder step j of a panel in one launch (the blocked QR's
// panel steps): workgroup 0 finishes the previous step's column (scales its
// vector and stores beta, from scr), then stores this step's tau[j] and
// beta / scale / flag in scr; workgroup b >= 1 applies H_j to column j + b.
// Every workgroup forms x(j+1:)^2's sum with the same workgroup reduction and
// v = x * scale with the same rounding, so all of them agree on the
// reflector.  Column j itself is written only by the next step's launch (or
// k_hh_finish), after every reader is done.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SK_NT) void k_hh_step(int64_t rows, int64_t j, double* __restrict__ W,
                                                   int64_t ld, double* __restrict__ tau,
                                                   double* __restrict__ scr, int has_prev) {
  __shared__ double red[SK_NT / 64];
  const double* x = W + j * ld;
  double acc = 0.0;
  for (int64_t i = j + 1 + threadIdx.x; i < rows; i += SK_NT) acc = fma(x[i], x[i], acc);
  const double xn2 = rbrp_block_sum(acc, red);
  const double alpha = x[j];
  double tj = 0.0, beta = 0.0, scal = 0.0;
  if (xn2 != 0.0) {
    const double nrm = sqrt(alpha * alpha + xn2);
    beta = (alpha >= 0.0) ? -nrm : nrm;
    scal = 1.0 / (alpha - beta);
    tj = (beta - alpha) / beta;
  }
  if (blockIdx.x == 0) {
    if (has_prev && scr[2] != 0.0) {
      double* y = W + (j - 1) * ld;
      const double sp = scr[0];
      for (int64_t i = j + threadIdx.x; i < rows; i += SK_NT) y[i] *= sp;
      if (threadIdx.x == 0) y[j - 1] = scr[1];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      tau[j] = tj;
      scr[0] = scal;
      scr[1] = beta;
      scr[2] = (xn2 != 0.0) ? 1.0 : 0.0;
    }
    return;
  }
  if (xn2 == 0.0) return;  // tau_j = 0: H_j is the identity
  double* w = W + (j + blockIdx.x) * ld;
  double dacc = 0.0;
  for (int64_t i = j + 1 + threadIdx.x; i < rows; i += SK_NT) dacc = fma(x[i] * scal, w[i], dacc);
  const double dot = rbrp_block_sum(dacc, red) + w[j];
  const double f = tj * dot;
  for (int64_t i = j + 1 + threadIdx.x; i < rows; i += SK_NT) w[i] -= f * (x[i] * scal);
  __syncthreads();
  if (threadIdx.x == 0) w[j] -= f;
}

// The last step's column: its vector scaled and beta stored (from scr).
__global__ __launch_bounds__(SK_NT) void k_hh_finish(int64_t rows, int64_t j, double* __restrict__ W,
                                                     int64_t ld, const double* __restrict__ scr) {
  if (scr[2] == 0.0) return;
  double* y = W + j * ld;
  const double sp = scr[0];
  for (int64_t i = j + 1 + threadIdx.x; i < rows; i += SK_NT) y[i] *= sp;
  if (threadIdx.x == 0) y[j] = scr[1];
}

hipError_t rbrp_launch_hh_step(hipStream_t st, int64_t rows, int64_t cend, int64_t j, double* W,
                               int64_t ld, double* tau, double* scr, int has_prev) {
  hipLaunchKernelGGL(k_hh_step, dim3((unsigned)(cend - j)), dim3(SK_NT), 0, st, rows, j, W, ld, tau, scr,
                     has_prev);
  return hipGetLastError();
}

hipError_t rbrp_launch_hh_finish(hipStream_t st, int64_t rows, int64_t j, double* W, int64_t ld,
                                 const double* scr) {
  hipLaunchKernelGGL(k_hh_finish, dim3(1), dim3(SK_NT), 0, st, rows, j, W, ld, scr);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LU with partial pivoting, step j of Y (nrows x ncols): pivot = first index of
// max |Y(i, j)| over i >= j (LAPACK idamax), swap rows j and pivot across all
// columns, ipiv[j] = pivot (0-based), scale Y(j+1:, j) by 1 / Y(j, j) (skipped
// for an exact zero pivot, as dgetf2 does).  One workgroup.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SK_NT) void k_lu_pivot(int64_t nrows, int64_t c0, int64_t ncols, int64_t j,
                                                    double* __restrict__ Y, int64_t ld,
                                                    int* __restrict__ ipiv) {
  __shared__ double rv[SK_NT];
  __shared__ int ri[SK_NT];
  __shared__ int s_p;
  const int tid = threadIdx.x;
  double bv = -1.0;
  int bi = 0x7fffffff;
  for (int64_t i = j + tid; i < nrows; i += SK_NT) {
    const double a = fabs(Y[i + j * ld]);
    if (a > bv) {
      bv = a;
      bi = (int)i;
    }
  }
  rv[tid] = bv;
  ri[tid] = bi;
  __syncthreads();
  for (int o = SK_NT / 2; o > 0; o >>= 1) {
    if (tid < o) {
      const double ov = rv[tid + o];
      const int oi = ri[tid + o];
      if (ov > rv[tid] || (ov == rv[tid] && oi < ri[tid])) {
        rv[tid] = ov;
        ri[tid] = oi;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    s_p = (ri[0] == 0x7fffffff) ? (int)j : ri[0];
    ipiv[j] = s_p;
  }
  __syncthreads();
  const int64_t p = s_p;
  if (p != j) {
    for (int64_t c = c0 + tid; c < ncols; c += SK_NT) {
      const double t0 = Y[j + c * ld];
      Y[j + c * ld] = Y[p + c * ld];
      Y[p + c * ld] = t0;
    }
    __syncthreads();
  }
  const double piv = Y[j + j * ld];
  if (piv != 0.0) {
    const double inv = 1.0 / piv;
    for (int64_t i = j + 1 + tid; i < nrows; i += SK_NT) Y[i + j * ld] *= inv;
  }
}

hipError_t rbrp_launch_lu_pivot(hipStream_t st, int64_t nrows, int64_t c0, int64_t ncols, int64_t j,
                                double* Y, int64_t ld, int* ipiv) {
  hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(SK_NT), 0, st, nrows, c0, ncols, j, Y, ld, ipiv);
  return hipGetLastError();
}

// The row interchanges ipiv[j0:j1) of a factored panel applied, in order, to
// every column c in [0, ncols) outside [j0, j1): one thread per column.
__global__ void k_lu_laswp(int64_t ncols, int64_t j0, int64_t j1, double* __restrict__ Y, int64_t ld,
                           const int* __restrict__ ipiv) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols || (c >= j0 && c < j1)) return;
  double* y = Y + c * ld;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t p = ipiv[j];
    if (p != j) {
      const double t = y[j];
      y[j] = y[p];
      y[p] = t;
    }
  }
}

hipError_t rbrp_launch_lu_laswp(hipStream_t st, int64_t ncols, int64_t j0, int64_t j1, double* Y,
                                int64_t ld, const int* ipiv) {
  if (ncols <= 0 || j1 <= j0) return hipSuccess;
  hipLaunchKernelGGL(k_lu_laswp, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, st, ncols, j0, j1, Y,
                     ld, ipiv);
  return hipGetLastError();
}

// Linv (jb x jb, ldi) <- L^{-1} for the unit lower-triangular jb x jb block L
// (strictly lower part at L, ld), jb <= 64: thread c forward-substitutes
// column c of the identity.
__global__ __launch_bounds__(64) void k_unit_lower_inv(int jb, const double* __restrict__ L, int64_t ld,
                                                       double* __restrict__ Linv, int64_t ldi) {
  __shared__ double Ls[64 * 64];  // L(i, r) at i + 64 r
  __shared__ double X[64 * 64];   // element i of column t of the inverse at t + 64 i
  const int t = threadIdx.x;
  for (int r = 0; r < jb; ++r)
    if (t < jb) Ls[t + 64 * r] = L[t + (int64_t)r * ld];
  __syncthreads();
  if (t < jb) {
    // (element-major layout: the 64 threads' columns sit in distinct banks)
    for (int i = 0; i < jb; ++i) {
      double xi;
      if (i < t) {
        xi = 0.0;
      } else if (i == t) {
        xi = 1.0;
      } else {
        double acc = 0.0;
        for (int r = t; r < i; ++r) acc = fma(Ls[i + 64 * r], X[t + 64 * r], acc);
        xi = -acc;
      }
      X[t + 64 * i] = xi;
    }
    for (int i = 0; i < jb; ++i) Linv[i + (int64_t)t * ldi] = X[t + 64 * i];
  }
}

hipError_t rbrp_launch_unit_lower_inv(hipStream_t st, int jb, const double* L, int64_t ld, double* Linv,
                                      int64_t ldi) {
  if (jb <= 0 || jb > 64) return jb <= 0 ? hipSuccess : hipErrorInvalidValue;
  hipLaunchKernelGGL(k_unit_lower_inv, dim3(1), dim3(64), 0, st, jb, L, ld, Linv, ldi);
  return hipGetLastError();
}

// Rank-1 update Y(j+1:, c) -= Y(j+1:, j) * Y(j, c) for c in (j, ncols); one
// workgroup per column.
__global__ __launch_bounds__(SK_NT) void k_lu_update(int64_t nrows, int64_t ncols, int64_t j,
                                                     double* __restrict__ Y, int64_t ld) {
  const int64_t c = j + 1 + blockIdx.x;
  if (c >= ncols) return;
  const double u = Y[j + c * ld];
  if (u == 0.0) return;
  const double* l = Y + j * ld;
  double* y = Y + c * ld;
  for (int64_t i = j + 1 + threadIdx.x; i < nrows; i += SK_NT) y[i] = fma(-l[i], u, y[i]);
}

hipError_t rbrp_launch_lu_update(hipStream_t st, int64_t nrows, int64_t ncols, int64_t j, double* Y,
                                 int64_t ld) {
  if (ncols - j - 1 <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lu_update, dim3((unsigned)(ncols - j - 1)), dim3(SK_NT), 0, st, nrows, ncols, j,
                     Y, ld);
  return hipGetLastError();
}
