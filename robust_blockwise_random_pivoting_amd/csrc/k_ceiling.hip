// Ceiling measurements of the part, for the benchmark's roofline figures
// (built into the test-hook library, not the product library): the rate of a
// streaming read, of a streaming copy, of an in-place sweep in the trailing
// update's access pattern, and of loops of independent FP64 / FP32
// matrix-core instructions, each timed with events over repeated launches.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/rbrp_testing.h"
#include "rbrp_common.h"

namespace {

// Streaming read: every lane sums its 16-byte units (nontemporal loads, four
// in flight), one store per workgroup keeps the loads live.
__global__ __launch_bounds__(256) void k_ceil_read(const v2d* __restrict__ p, int64_t n2,
                                                   double* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  v2d a0 = {0.0, 0.0}, a1 = {0.0, 0.0};
  for (; q + 3 * stride < n2; q += 4 * stride) {
    const v2d x0 = __builtin_nontemporal_load(p + q), x1 = __builtin_nontemporal_load(p + q + stride),
              x2 = __builtin_nontemporal_load(p + q + 2 * stride),
              x3 = __builtin_nontemporal_load(p + q + 3 * stride);
    a0 += x0 + x2;
    a1 += x1 + x3;
  }
  for (; q < n2; q += stride) a0 += __builtin_nontemporal_load(p + q);
  const v2d a = a0 + a1;
  if (threadIdx.x == 0) out[blockIdx.x] = a.x + a.y;
}

// Streaming copy (nontemporal loads and stores).
__global__ __launch_bounds__(256) void k_ceil_copy(const v2d* __restrict__ s, v2d* __restrict__ d,
                                                   int64_t n2) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n2; q += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(s + q), d + q);
}

// Independent FP64 matrix-core chains: each wave issues `iters` x 8
// v_mfma_f64_16x16x4f64 on eight accumulators (2 * 16 * 16 * 4 flops each).
__global__ __launch_bounds__(256) void k_ceil_mfma64(int iters, double seed, double* __restrict__ out) {
  const double a = seed + threadIdx.x * 1e-9, b = 1.0 - threadIdx.x * 1e-9;
  v4d c[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) c[u] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[u], 0, 0, 0);
  }
  double t = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) t += c[u][0] + c[u][1] + c[u][2] + c[u][3];
  if (t == 12345.678) out[blockIdx.x] = t;  // never true; keeps the chains live
}

// In-place read-modify-write sweep of a column-major rows x cols matrix in
// the trailing update's access pattern: a workgroup takes a 128-column tile
// (32 columns per wave) over T = 8 row tiles of 64 rows, each wave moving
// 16-row x 4-column pieces per load / store instruction (the matrix-core
// output layout of the update), two workgroups per compute unit.
__global__ __launch_bounds__(256, 2) void k_ceil_sweep(double* __restrict__ C, int64_t rows,
                                                       int64_t cols, int ncol) {
  const int T = 8;
  const int64_t bx = blockIdx.x % ncol, by = blockIdx.x / ncol;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int64_t cw = bx * 128 + 32 * wid;
  if (cw >= cols) return;
  for (int t = 0; t < T; ++t) {
    const int64_t r0 = (by * T + t) * 64;
    if (r0 >= rows) break;
    double v[32];
#pragma unroll
    for (int sb = 0; sb < 4; ++sb)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t c = min(cw + 16 * (q >> 2) + 4 * (q & 3) + lk, cols - 1);
        v[8 * sb + q] = C[r0 + 16 * sb + li + c * rows];
      }
#pragma unroll
    for (int sb = 0; sb < 4; ++sb)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t c = cw + 16 * (q >> 2) + 4 * (q & 3) + lk;
        if (c < cols) C[r0 + 16 * sb + li + c * rows] = fma(v[8 * sb + q], 0.5, 1.0);
      }
  }
}

typedef float v4f_c __attribute__((ext_vector_type(4)));

// The same with v_mfma_f32_16x16x4f32 on sixteen accumulators (the fastest
// of 8 / 16 chains measured).
__global__ __launch_bounds__(256) void k_ceil_mfma32(int iters, float seed, double* __restrict__ out) {
  const float a = seed + threadIdx.x * 1e-6f, b = 1.0f - threadIdx.x * 1e-6f;
  v4f_c c[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) c[u] = (v4f_c){0.0f, 0.0f, 0.0f, 0.0f};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 16; ++u) c[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c[u], 0, 0, 0);
  }
  float t = 0.0f;
#pragma unroll
  for (int u = 0; u < 16; ++u) t += c[u][0] + c[u][1] + c[u][2] + c[u][3];
  if (t == 12345.678f) out[blockIdx.x] = t;
}

}  // namespace

extern "C" int rbrp_test_ceiling(void* stream, int kind, int64_t bytes, int reps, double* ms_per_launch,
                                 double* rate) {
  if (kind < 0 || kind > 4) return -2;
  if (bytes < 1 << 20) return -3;
  if (reps < 1) return -4;
  if (!ms_per_launch || !rate) return -5;
  hipStream_t st = (hipStream_t)stream;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return (int)hipErrorUnknown;
  const int64_t n2 = bytes / 16;
  // sweep: 50048 rows (a multiple of 64), as many columns as `bytes` holds
  const int64_t srows = 50048, scols = std::max<int64_t>(128, bytes / (8 * srows));
  const int sncol = (int)((scols + 127) / 128), snrg = (int)((srows / 64 + 7) / 8);
  void *a = nullptr, *b = nullptr;
  double* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&out, sizeof(double) * 65536);
  if (e == hipSuccess && kind <= 1) e = hipMalloc(&a, (size_t)n2 * 16);
  if (e == hipSuccess && kind == 1) e = hipMalloc(&b, (size_t)n2 * 16);
  if (e == hipSuccess && kind <= 1) e = hipMemsetAsync(a, 0, (size_t)n2 * 16, st);
  if (e == hipSuccess && kind == 4) e = hipMalloc(&a, (size_t)(srows * scols * 8));
  if (e == hipSuccess && kind == 4) e = hipMemsetAsync(a, 0, (size_t)(srows * scols * 8), st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  // (launches of 1-3 ms: longer all-matrix-core launches run at lower clocks)
  const int grid_mem = cus * 8, grid_mfma = cus * 8, iters = 1024;
  auto launch = [&]() {
    if (kind == 0)
      hipLaunchKernelGGL(k_ceil_read, dim3(grid_mem), dim3(256), 0, st, (const v2d*)a, n2, out);
    else if (kind == 1)
      hipLaunchKernelGGL(k_ceil_copy, dim3(grid_mem), dim3(256), 0, st, (const v2d*)a, (v2d*)b, n2);
    else if (kind == 4)
      hipLaunchKernelGGL(k_ceil_sweep, dim3(sncol * snrg), dim3(256), 0, st, (double*)a, srows, scols,
                         sncol);
    else if (kind == 2)
      hipLaunchKernelGGL(k_ceil_mfma64, dim3(grid_mfma), dim3(256), 0, st, iters, 0.5, out);
    else
      hipLaunchKernelGGL(k_ceil_mfma32, dim3(grid_mfma), dim3(256), 0, st, iters, 0.5f, out);
  };
  for (int w = 0; w < 2 && e == hipSuccess; ++w) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (out) (void)hipFree(out);
  if (e != hipSuccess) return (int)e;
  const double t = ms / reps;
  *ms_per_launch = t;
  if (kind == 0)
    *rate = (double)n2 * 16 / (t * 1e-3) / 1e9;  // GB/s read
  else if (kind == 1)
    *rate = (double)n2 * 32 / (t * 1e-3) / 1e9;  // GB/s read + written
  else if (kind == 4)
    *rate = (double)srows * scols * 16 / (t * 1e-3) / 1e9;  // GB/s read + written
  else
    *rate = (double)grid_mfma * 4 * iters * (kind == 2 ? 8 : 16) * (2.0 * 16 * 16 * 4) / (t * 1e-3) /
            1e12;  // TFLOP/s
  return 0;
}
