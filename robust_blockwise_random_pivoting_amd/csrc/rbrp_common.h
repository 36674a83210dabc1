// Shared device/host helpers for the RBRP HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RBRP_WAVE 64

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Counter-based: block q of the stream is
// Philox(counter = base + 1 + q, key); draw t is word t % 4 of block t / 4.
// This is the word order and counter convention of numpy.random.Philox, so a
// NumPy Generator built on Philox(key=seed, counter=base) consumes the same
// stream (checked by the tests against NumPy and the Random123 KAT vectors).
// ---------------------------------------------------------------------------
struct Philox4x64Out {
  uint64_t v[4];
};

__host__ __device__ inline uint64_t rbrp_mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
#endif
}

__host__ __device__ inline Philox4x64Out rbrp_philox4x64_10(const uint64_t ctr_in[4],
                                                           uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t lo0 = M0 * c0, hi0 = rbrp_mulhi64(M0, c0);
    const uint64_t lo1 = M1 * c2, hi1 = rbrp_mulhi64(M1, c2);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  Philox4x64Out o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// 64-bit word number `t` of the stream whose counter starts at `base` (256-bit,
// little-endian words): block counter = base + 1 + t / 4 with carry.
__host__ __device__ inline uint64_t rbrp_philox_word(const uint64_t base[4], uint64_t k0,
                                                     uint64_t k1, uint64_t t) {
  uint64_t ctr[4];
  uint64_t add = (t >> 2) + 1;
  uint64_t carry = 0;
  // add `add` into word 0 and propagate the carry (add >= 1 and fits in 63 bits)
  ctr[0] = base[0] + add;
  carry = (ctr[0] < base[0]) ? 1 : 0;
  ctr[1] = base[1] + carry;
  carry = (carry && ctr[1] == 0) ? 1 : 0;
  ctr[2] = base[2] + carry;
  carry = (carry && ctr[2] == 0) ? 1 : 0;
  ctr[3] = base[3] + carry;
  Philox4x64Out o = rbrp_philox4x64_10(ctr, k0, k1);
  return o.v[t & 3];
}

// NumPy's uint64 -> [0, 1) double conversion: top 53 bits times 2^-53.
__host__ __device__ inline double rbrp_u64_to_unit(uint64_t x) {
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Wave / workgroup reductions (fixed order, so results are deterministic)
// ---------------------------------------------------------------------------
__device__ inline double rbrp_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, RBRP_WAVE);
  return v;
}

__device__ inline int rbrp_wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, RBRP_WAVE);
  return v;
}

// Sum over the workgroup; `red` must hold blockDim.x / 64 doubles.  Every
// thread returns the total.  Wave partials are added in wave order.
__device__ inline double rbrp_block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = rbrp_wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += red[w];
  return t;
}

__device__ inline int rbrp_block_sum_int(int v, int* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = rbrp_wave_sum_int(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  int t = 0;
  for (int w = 0; w < nw; ++w) t += red[w];
  return t;
}

// XCD-aware workgroup order.  The hardware hands workgroups to the 8 XCDs
// (each with its own L2) round-robin by linear index.  A kernel launched on a
// 1-D grid of rbrp_xcd_grid(W) workgroups maps workgroup L to the logical task
// rbrp_xcd_task(L, W) = (L % 8) * ceil(W / 8) + L / 8, so each XCD runs one
// contiguous range of tasks (and tasks that share an operand, numbered
// consecutively, share that XCD's L2); tasks >= W are empty.
static inline int64_t rbrp_xcd_grid(int64_t W) { return (W + 7) / 8 * 8; }
#if defined(__HIPCC__)
__device__ inline int64_t rbrp_xcd_task(int64_t L, int64_t W) {
  return (L % 8) * ((W + 7) / 8) + L / 8;
}
#endif

// Per-block device state shared by the kernels of one RBRP block step.
struct RbrpBlockState {
  int s;            // number of sampled columns
  int bt;           // columns kept by the robust filter
  int nswap;        // |Jin| = |Jout|
  int status;       // 0 ok; nonzero = numerical failure in the block
  unsigned long long draws;  // cumulative uniforms consumed from the stream
  double resid;     // sum of residual column norms^2 after the block
  double err;       // resid / N0
  double N0;        // ||A||_F^2
  int nnz;          // nonzero weights seen by the sampler
  int seq;          // publication number of the host copy (k_resid_total)
};
