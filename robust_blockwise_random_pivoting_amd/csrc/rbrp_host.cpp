// Host orchestration of the RBRP column ID and the C ABI (include/rbrp.h).
//
// One handle = one stream + one cached device workspace.  The block loop
// follows RBRP() of gpu/rbrp.cu:333-503 in right-looking form: per block the
// host launches the sampler, the re-orthogonalisation, the panel (Gram +
// pivoted Cholesky + filter + CholQR2), the permutation and the trailing
// projection/update, then reads one small state record (kept count, residual)
// back to decide whether to continue.  All sizes that depend on the block's
// data (sampled count s, kept count bt) are read by the kernels from device
// memory, so no other host round trip happens inside a block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rbrp.h"
#include "rbrp_handle.h"
#include "rbrp_kernels.h"

#define HIPCHK(x)                               \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) return (int)e_;       \
  } while (0)

namespace {

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct Carver {
  size_t off = 0;
  template <typename T>
  size_t take(int64_t count) {
    size_t o = off;
    off += round_up((int64_t)(count * sizeof(T)), 256);
    return o;
  }
};

}  // namespace

namespace {

// Byte offsets of the loop's buffers in one arena (every buffer 256-byte aligned).
struct WsLayout {
  size_t oAw, oAw32, oQ64, oVd, oRcol, oR64, oR, odiags, ow, ocdf, oG, oG2, oRb, oRinv, oR2, oR2inv, oCbuf, oQbuf, oP,
      onp, oWt, oJ, oJk, oJin, oJout, oJkpos, oDinv, ost;
  size_t need;
  int64_t Pcap;
};

// Split-K partials of the full-size arena: 64 splits of a 128-row product
// over max(n, kmax) columns.
int64_t full_pcap(int64_t n, int64_t kmax) {
  return 64 * RBRP_MAX_B * std::max<int64_t>(std::max(n, kmax), RBRP_MAX_B);
}

// `mixed`: the FP32 working matrix, the FP64 skeleton basis and the buffers of
// the FP64 recompute; `w64`: the FP64 working matrix (the FP64 loop, and the
// mixed loop's FP64 first blocks); `vd`: the FP64 block buffer Vd (the mixed
// loop's sampled block, the FP64 loop's first sampled block when that block
// reads A directly).
WsLayout layout_workspace(int64_t mp, int64_t n, int64_t kmax, int64_t b, bool mixed, bool w64,
                          int64_t Pcap, bool vd) {
  const int64_t ldr = std::max<int64_t>(kmax, 1);
  WsLayout L;
  Carver c;
  L.oAw = w64 ? c.take<double>(mp * n) : 0;
  L.oAw32 = mixed ? c.take<float>(mp * n) : 0;
  L.oQ64 = mixed ? c.take<double>(mp * ldr) : 0;
  L.oVd = (mixed || vd) ? c.take<double>(mp * b) : 0;
  // mixed loop: per-chunk totals and the total of the FP64 residual recompute
  L.oRcol = mixed ? c.take<double>(n) : 0;
  L.oR64 = mixed ? c.take<double>(1) : 0;
  L.oR = c.take<double>(ldr * n);
  L.ost = c.take<RbrpBlockState>(2);  // the block state and its published snapshot (after R:
                                      // one fill kernel clears both)
  L.odiags = c.take<double>(n);
  L.ow = c.take<double>(n);
  L.ocdf = c.take<double>(n);
  L.oG = c.take<double>(b * b);
  L.oG2 = c.take<double>(b * b);
  L.oRb = c.take<double>(b * b);
  L.oRinv = c.take<double>(b * b);
  L.oR2 = c.take<double>(b * b);
  L.oR2inv = c.take<double>(b * b);
  L.oCbuf = c.take<double>(ldr * b);
  L.oQbuf = c.take<double>(mp * b);
  L.oP = c.take<double>(Pcap);
  L.onp = c.take<double>(((mp + 63) / 64) * n);  // 64-row (FP64) or 128-row (FP32) tiles
  L.oWt = c.take<double>(b * n);                  // FP32 in the mixed loop (over-sized)
  L.oJ = c.take<int>(b);
  L.oJk = c.take<int>(b);
  L.oJin = c.take<int>(b);
  L.oJout = c.take<int>(b);
  L.oJkpos = c.take<int>(b);
  L.oDinv = c.take<double>(4 * ((kmax + 255) / 256) * 4096);  // diagonal-block inverses
  L.need = c.off;
  L.Pcap = Pcap;
  return L;
}

// Points the workspace's buffers into the arena at `base`.
void carve_workspace(RbrpWorkspace& ws, char* base, const WsLayout& L, bool mixed, bool w64,
                     bool vd) {
  ws.Aw = w64 ? (double*)(base + L.oAw) : nullptr;
  ws.Aw32 = mixed ? (float*)(base + L.oAw32) : nullptr;
  ws.Q64 = mixed ? (double*)(base + L.oQ64) : nullptr;
  ws.Vd = (mixed || vd) ? (double*)(base + L.oVd) : nullptr;
  ws.rcol = mixed ? (double*)(base + L.oRcol) : nullptr;
  ws.r64 = mixed ? (double*)(base + L.oR64) : nullptr;
  ws.R = (double*)(base + L.oR);
  ws.diags = (double*)(base + L.odiags);
  ws.w = (double*)(base + L.ow);
  ws.cdf = (double*)(base + L.ocdf);
  ws.G = (double*)(base + L.oG);
  ws.G2 = (double*)(base + L.oG2);
  ws.Rb = (double*)(base + L.oRb);
  ws.Rinv = (double*)(base + L.oRinv);
  ws.R2 = (double*)(base + L.oR2);
  ws.R2inv = (double*)(base + L.oR2inv);
  ws.Cbuf = (double*)(base + L.oCbuf);
  ws.Qbuf = (double*)(base + L.oQbuf);
  ws.P = (double*)(base + L.oP);
  ws.normpart = (double*)(base + L.onp);
  ws.Wt = (double*)(base + L.oWt);  // (one of the two per block: FP64 or FP32 storage)
  ws.Wt32 = mixed ? (float*)(base + L.oWt) : nullptr;
  ws.J = (int*)(base + L.oJ);
  ws.Jk = (int*)(base + L.oJk);
  ws.Jin = (int*)(base + L.oJin);
  ws.Jout = (int*)(base + L.oJout);
  ws.Jkpos = (int*)(base + L.oJkpos);
  ws.Dinv = (double*)(base + L.oDinv);
  ws.st = (RbrpBlockState*)(base + L.ost);
  ws.st_snap = ws.st + 1;
  ws.Pcap = L.Pcap;
}

// The handle's cached arena, grown on demand.
int ensure_workspace(rbrp_handle_t h, int64_t mp, int64_t n, int64_t kmax, int64_t b, bool mixed,
                     bool w64, bool vd) {
  const WsLayout L = layout_workspace(mp, n, kmax, b, mixed, w64, full_pcap(n, kmax), vd);
  RbrpWorkspace& ws = h->ws;
  if (ws.arena_bytes < L.need) {
    if (ws.arena) {
      (void)hipFree(ws.arena);
      ws.arena = nullptr;
      ws.arena_bytes = 0;
    }
    hipError_t e = hipMalloc(&ws.arena, L.need);
    if (e != hipSuccess) {
      ws.arena = nullptr;
      return RBRP_ERR_ALLOC;
    }
    ws.arena_bytes = L.need;
  }
  carve_workspace(ws, (char*)ws.arena, L, mixed, w64, vd);
  return 0;
}

// ---- static scratch of the no-allocation entry point -----------------------
// One array per device, loaded with the library's code object and held until
// the process ends.  Host code never uses the array's name as an address: the
// device address of each GPU's copy is looked up once per device (rbrp_create,
// with that device current) and kept in g_scratch.
static __device__ __align__(256) unsigned char scratch[RBRP_NOALLOC_SCRATCH_BYTES];

struct ScratchSlot {
  std::atomic<int> state{0};  // 0 not looked up, 1 looking up, 2 done
  void* ptr = nullptr;
  hipError_t err = hipSuccess;
};
ScratchSlot g_scratch[RBRP_MAX_DEVICES];

// Looks up the current device's scratch address (once per device).
hipError_t scratch_setup(int dev) {
  if (dev < 0 || dev >= RBRP_MAX_DEVICES) return hipErrorInvalidDevice;
  ScratchSlot& sl = g_scratch[dev];
  int expect = 0;
  if (sl.state.compare_exchange_strong(expect, 1, std::memory_order_acq_rel)) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    sl.ptr = e == hipSuccess ? ptr : nullptr;
    sl.err = e;
    sl.state.store(2, std::memory_order_release);
    return e;
  }
  while (sl.state.load(std::memory_order_acquire) != 2) {
  }
  return sl.err;
}

// The device's scratch address, or null (with *err set) if the lookup failed
// or was never made on this device.
void* scratch_ptr(int dev, hipError_t* err) {
  if (dev < 0 || dev >= RBRP_MAX_DEVICES) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  ScratchSlot& sl = g_scratch[dev];
  if (sl.state.load(std::memory_order_acquire) != 2) {
    *err = hipErrorNotInitialized;
    return nullptr;
  }
  *err = sl.err;
  return sl.ptr;
}

// Carves the loop's workspace out of the device's static scratch: with the
// arena path's split-K partials when they fit (same launch shapes, so the
// same results as rbrp_colid_d), else with partials for at most 16 splits.
int static_workspace(rbrp_handle_t h, int64_t mp, int64_t n, int64_t kmax, int64_t b, bool mixed,
                     bool w64, bool vd, RbrpWorkspace* ws) {
  hipError_t e = hipSuccess;
  void* base = scratch_ptr(h->device, &e);
  if (!base) return e != hipSuccess ? (int)e : (int)hipErrorNotInitialized;
  WsLayout L = layout_workspace(mp, n, kmax, b, mixed, w64, full_pcap(n, kmax), vd);
  if (L.need > RBRP_NOALLOC_SCRATCH_BYTES)
    L = layout_workspace(mp, n, kmax, b, mixed, w64, 16 * b * std::max<int64_t>(std::max(n, kmax), b),
                         vd);
  if (L.need > RBRP_NOALLOC_SCRATCH_BYTES) return RBRP_ERR_SCRATCH;
  *ws = RbrpWorkspace();
  carve_workspace(*ws, (char*)base, L, mixed, w64, vd);
  return 0;
}

// Split-K shape for a TN product of Mmax x Nmax over K rows: aim for about
// 1024 workgroups, keep the partials within the workspace.
// Split count and chunk length of a split-K TN launch: as many splits as fill
// `resident` workgroups (the kernel's whole-device residency: one round, no
// tail of a few late workgroups) or about 1024 workgroups when that is unknown.
void choose_split(int64_t Mmax, int64_t Nmax, int64_t K, int64_t ldp, int64_t Pcap, int* nsplit,
                  int64_t* kchunk, int64_t align = 16, int64_t tile_n = 64, int resident = 0,
                  int64_t tile_m = 64) {
  const int64_t tiles = ((Mmax + tile_m - 1) / tile_m) * ((Nmax + tile_n - 1) / tile_n);
  const int64_t target = resident > 0 ? resident : 1024;
  int64_t ns = std::max<int64_t>(1, target / std::max<int64_t>(tiles, 1));
  ns = std::min<int64_t>(ns, tiles <= 4 ? 256 : 128);
  ns = std::min<int64_t>(ns, std::max<int64_t>(1, Pcap / std::max<int64_t>(ldp * Nmax, 1)));
  ns = std::min<int64_t>(ns, std::max<int64_t>(1, K / 64));
  int64_t kc = round_up((K + ns - 1) / ns, align);
  ns = (K + kc - 1) / kc;
  *nsplit = (int)std::max<int64_t>(ns, 1);
  *kchunk = kc;
}

// Per-phase event timing: events are recorded around launches (or handed to a
// launch that times its own dispatch) and read when the ID ends, so the blocks
// run without event waits in between.
struct PhaseTimer {
  rbrp_handle_t h;
  struct Rec {
    int phase;
    int e0, e1;
    int launches;
    double flops, bytes;
    bool by_kernel;  // events recorded by the kernel's dispatch
  };
  std::vector<Rec> recs;
  int next = 0;
  bool enabled;  // false: no events at all (the no-allocation entry point)
  explicit PhaseTimer(rbrp_handle_t hh, bool on = true) : h(hh), enabled(on) {}
  hipEvent_t ev(int idx) {
    while ((int)h->events.size() <= idx) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      h->events.push_back(e);
    }
    return h->events[idx];
  }
  int open(int phase, bool by_kernel) {
    if (!enabled || !h->timing) return -1;
    if (h->timing == 2 && phase != RBRP_PHASE_PROJECT && phase != RBRP_PHASE_UPDATE) return -1;
    Rec r{phase, next, next + 1, 0, 0.0, 0.0, by_kernel};
    next += 2;
    if (!ev(r.e0) || !ev(r.e1)) return -1;
    recs.push_back(r);
    return (int)recs.size() - 1;
  }
  // A phase of several launches: marker events recorded around them.
  int begin(int phase) {
    const int id = open(phase, false);
    if (id >= 0) (void)hipEventRecord(h->events[recs[id].e0], h->stream);
    return id;
  }
  // A phase that is one launch: its events are handed to the launch and timed
  // by the dispatch itself (*e0 / *e1 stay null when the phase is not timed).
  int begin_kernel(int phase, hipEvent_t* e0, hipEvent_t* e1) {
    const int id = open(phase, true);
    *e0 = id >= 0 ? h->events[recs[id].e0] : nullptr;
    *e1 = id >= 0 ? h->events[recs[id].e1] : nullptr;
    return id;
  }
  void end(int id, int launches, double flops = 0.0, double bytes = 0.0) {
    if (id < 0) return;
    Rec& r = recs[id];
    r.launches = launches;
    r.flops = flops;
    r.bytes = bytes;
    if (!r.by_kernel) (void)hipEventRecord(h->events[r.e1], h->stream);
  }
  // Set the work figures of a record once the block's sizes are known.
  void set_work(int id, double flops, double bytes) {
    if (id < 0) return;
    recs[id].flops = flops;
    recs[id].bytes = bytes;
  }
  void flush() {  // waits for each record's end event
    for (const Rec& r : recs) {
      float ms = 0.f;
      if (hipEventSynchronize(h->events[r.e1]) == hipSuccess &&
          hipEventElapsedTime(&ms, h->events[r.e0], h->events[r.e1]) == hipSuccess)
        h->stats.ms[r.phase] += ms;
      h->stats.launches[r.phase] += r.launches;
      h->stats.flops[r.phase] += r.flops;
      h->stats.bytes[r.phase] += r.bytes;
    }
    recs.clear();
    next = 0;
  }
};

// Wait until the block's last kernel has published state number `seq` into
// the pinned host copy (k_resid_total): spin on the sequence number, polling
// the stream now and then so that an error or a stream that finished without
// the publication ends the wait.  In the latter case the record is copied from
// the device snapshot the same kernel wrote (`snap`, which the next block's
// queued kernels do not touch); a snapshot of another block is an error.
hipError_t wait_published(rbrp_handle_t h, hipStream_t st, const RbrpBlockState* snap, int seq) {
  for (uint64_t it = 1;; ++it) {
    if (__atomic_load_n(&h->host_state->seq, __ATOMIC_ACQUIRE) == seq) return hipSuccess;
    if ((it & 1023) == 0) {
      const hipError_t q = hipStreamQuery(st);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) return q;
      if (__atomic_load_n(&h->host_state->seq, __ATOMIC_ACQUIRE) == seq) return hipSuccess;
      RbrpBlockState rec;
      const hipError_t e = hipMemcpy(&rec, snap, sizeof(RbrpBlockState), hipMemcpyDeviceToHost);
      if (e != hipSuccess) return e;
      if (rec.seq != seq) return hipErrorUnknown;
      *h->host_state = rec;
      return hipSuccess;
    }
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
  }
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// X (k x nc, ldx) <- R^{-1} X for the k x k upper-triangular R (ldr).  With a
// scratch `dinv` (4 * ceil(k / 256) x 4096 doubles) the 64-row diagonal
// blocks are inverted in one launch and every column tile then runs the whole
// backward blocked solve in one more (k_trsm_cols).  Without it: 256-row
// super-blocks, each solved with 64-row diagonal blocks by substitution
// (k_trsm_diag, one thread per column) and small updates, followed by one
// large update of every row above the super-block (k_gemm_nn).
hipError_t rbrp_trsm_upper_left(hipStream_t st, int64_t k, int64_t nc, double* R, int64_t ldr,
                                double* X, int64_t ldx, int* launches, double* dinv) {
  hipError_t e;
  if (dinv && k > 0 && nc > 0) {
    // the diagonal-block inverses, then every block of every column tile in
    // one launch
    e = rbrp_launch_upper_inv_blocks(st, k, R, ldr, dinv);
    if (e != hipSuccess) return e;
    e = rbrp_launch_trsm_cols(st, k, nc, R, ldr, dinv, X, ldx);
    if (e != hipSuccess) return e;
    if (launches) *launches += 2;
    return hipSuccess;
  }
  for (int64_t s1 = k; s1 > 0; s1 -= 256) {
    const int64_t s0 = std::max<int64_t>(0, s1 - 256);
    for (int64_t r1 = s1; r1 > s0; r1 -= 64) {
      const int64_t r0 = std::max<int64_t>(s0, r1 - 64);
      if (r1 < s1) {
        e = rbrp_launch_gemm_nn(st, dim_static(r1 - r0), r1 - r0, dim_static(nc), nc,
                                dim_static(s1 - r1), mat_off(R + r0, ldr, r1), mat(X + r1, ldx),
                                mat(X + r0, ldx), GEMM_SUB, nullptr, 0);
        if (e != hipSuccess) return e;
        if (launches) ++*launches;
      }
      e = rbrp_launch_trsm_diag(st, r1 - r0, nc, R + r0 + r0 * ldr, ldr, X + r0, ldx);
      if (e != hipSuccess) return e;
      if (launches) ++*launches;
    }
    if (s0 > 0) {
      e = rbrp_launch_gemm_nn(st, dim_static(s0), s0, dim_static(nc), nc, dim_static(s1 - s0),
                              mat_off(R, ldr, s0), mat(X + s0, ldx), mat(X, ldx), GEMM_SUB, nullptr,
                              0);
      if (e != hipSuccess) return e;
      if (launches) ++*launches;
    }
  }
  return hipSuccess;
}

hipError_t rbrp_sketch_arena(rbrp_handle_t h, size_t bytes, void** out) {
  if (h->sk_bytes < bytes) {
    if (h->sk_arena) {
      hipError_t e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) return e;
      (void)hipFree(h->sk_arena);
      h->sk_arena = nullptr;
      h->sk_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->sk_arena, bytes);
    if (e != hipSuccess) {
      h->sk_arena = nullptr;
      return e;
    }
    h->sk_bytes = bytes;
  }
  *out = h->sk_arena;
  return hipSuccess;
}

extern "C" {

const char* rbrp_version(void) { return "rbrp-mi355x 0.2 (gfx950, FP64 and FP32 MFMA)"; }

void rbrp_opts_default(rbrp_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->b = 64;
  o->tol = 0.0;
  o->precision = RBRP_PREC_FP64;
  o->filter = RBRP_FILTER_SAMPLED;
}

void rbrp_philox4x64_10_host(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  Philox4x64Out o = rbrp_philox4x64_10(ctr, key[0], key[1]);
  for (int q = 0; q < 4; ++q) out[q] = o.v[q];
}

int rbrp_create(rbrp_handle_t* handle, void* stream) {
  if (!handle) return -1;
  *handle = nullptr;
  rbrp_handle_t h = new (std::nothrow) rbrp_handle_s();
  if (!h) return RBRP_ERR_ALLOC;
  h->stream = (hipStream_t)stream;
  std::memset(&h->stats, 0, sizeof(h->stats));
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    delete h;
    return (int)e;
  }
  e = hipHostMalloc((void**)&h->host_state, sizeof(RbrpBlockState),
                    hipHostMallocMapped | hipHostMallocCoherent);
  if (e != hipSuccess) {
    delete h;
    return (int)e;
  }
  e = hipHostMalloc((void**)&h->host_scalar, 16 * sizeof(int), hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipHostFree(h->host_state);
    delete h;
    return (int)e;
  }
  // the no-allocation entry point's scratch address on this device (a failed
  // lookup is reported by that entry point, not here)
  (void)scratch_setup(h->device);
  h->tn64_resident = rbrp_gemm_tn64_resident();
  h->tn32_resident = rbrp_gemm_tn32_resident(0);
  h->tn32w_resident = rbrp_gemm_tn32_resident(1);
  h->nn64_resident = rbrp_gemm_nn64_resident();
  *handle = h;
  return 0;
}

int rbrp_destroy(rbrp_handle_t h) {
  if (!h) return -1;
  rbrp_batched_release(h);
  if (h->ws.arena) (void)hipFree(h->ws.arena);
  if (h->sk_arena) (void)hipFree(h->sk_arena);
  if (h->host_state) (void)hipHostFree(h->host_state);
  if (h->host_scalar) (void)hipHostFree(h->host_scalar);
  for (hipEvent_t e : h->events) (void)hipEventDestroy(e);
  delete h;
  return 0;
}

int rbrp_set_stream(rbrp_handle_t h, void* stream) {
  if (!h) return -1;
  h->stream = (hipStream_t)stream;
  return 0;
}

int rbrp_set_timing(rbrp_handle_t h, int enable) {
  if (!h) return -1;
  if (enable < 0 || enable > 2) return -2;
  h->timing = enable;
  return 0;
}

int rbrp_get_phase_stats(rbrp_handle_t h, rbrp_phase_stats_t* out) {
  if (!h) return -1;
  if (!out) return -2;
  *out = h->stats;
  return 0;
}

int rbrp_reset_phase_stats(rbrp_handle_t h) {
  if (!h) return -1;
  std::memset(&h->stats, 0, sizeof(h->stats));
  return 0;
}

int rbrp_release_workspace(rbrp_handle_t h) {
  if (!h) return -1;
  if (h->ws.arena) {
    HIPCHK(hipStreamSynchronize(h->stream));
    (void)hipFree(h->ws.arena);
  }
  h->ws.arena = nullptr;
  h->ws.arena_bytes = 0;
  rbrp_batched_release(h);
  return 0;
}

int rbrp_uniforms_d(rbrp_handle_t h, uint64_t seed, uint64_t offset, uint64_t draw0, int64_t count,
                    double* dOut) {
  if (!h) return -1;
  if (count < 0) return -5;
  if (count > 0 && !dOut) return -6;
  const uint64_t ctr[4] = {offset, 0, 0, 0};
  HIPCHK(rbrp_launch_uniforms(h->stream, seed, 0, ctr, draw0, count, dOut));
  return 0;
}

}  // extern "C"

// The block loop for both precisions (also the per-problem path of the
// batched entry point).
int rbrp_colid_impl(rbrp_handle_t h, const void* dA, bool a_f32, int64_t m, int64_t n, int64_t lda,
                    const rbrp_opts_t* opts, rbrp_result_t* res);

namespace {

// The block loop for both precisions.  In the FP64 loop the working matrix Aw
// (FP64) holds the skeleton basis Q in columns [0, i) and the residual in
// [i, n).  In the mixed loop Aw32 (FP32) holds the residual and an FP32 copy
// of Q; the FP64 Q lives in Q64 and the sampled block is re-orthogonalised and
// factorised in the FP64 buffer Vd.
int colid_impl(rbrp_handle_t h, const void* dA, bool a_f32, int64_t m, int64_t n, int64_t lda,
               const rbrp_opts_t* opts, rbrp_result_t* res, bool noalloc = false) {
  const double t_start = now_s();
  if (!h) return -1;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (lda < std::max<int64_t>(m, 1)) return -5;
  if (!opts) return -6;
  if (!res) return -7;
  if (m > 0 && n > 0 && !dA) return -2;
  if (opts->b < 1 || opts->b > RBRP_MAX_B) return -6;
  if (!(opts->tol >= 0.0) || !std::isfinite(opts->tol)) return -6;
  if (opts->max_rank < 0) return -6;
  if (opts->precision != RBRP_PREC_FP64 && opts->precision != RBRP_PREC_FP32) return -6;
  if (opts->filter != RBRP_FILTER_SAMPLED && opts->filter != RBRP_FILTER_NOMINAL) return -6;
  if (n > 0 && !res->perm) return -7;
  if (m > 0x7fffffffLL || n > 0x7fffffffLL) return -3;
  const bool mixed = opts->precision == RBRP_PREC_FP32;

  const int64_t p = std::min(m, n);
  const int64_t kmax = opts->max_rank > 0 ? std::min<int64_t>(opts->max_rank, p) : p;
  if (res->X && res->ldx < std::max<int64_t>(kmax, 1)) return -7;
  if (res->trace_cap > 0 && (!res->ranks || !res->error || !res->t_ranks)) return -7;
  res->k = 0;
  res->nblocks = 0;
  res->err_final = 0.0;
  res->draws = 0;
  hipStream_t st = h->stream;

  if (n == 0) {
    res->t_total = now_s() - t_start;
    return 0;
  }
  if (m == 0) {
    HIPCHK(rbrp_launch_iota(st, res->perm, n));
    HIPCHK(hipStreamSynchronize(st));
    res->t_total = now_s() - t_start;
    return 0;
  }

  const int b = (int)opts->b;
  const int64_t mp = round_up(m, 64);
  const int64_t ldr = std::max<int64_t>(kmax, 1);
  // Storage schedule of the mixed loop.  An FP32 update of a residual of
  // relative size rho adds rounding noise of about 8 u32 rho ||A||_F (u32 =
  // 2^-24; at most 6.4 measured on the exp-decay matrices), and the residual
  // shrinks by a factor of about two per block, so the noise of all FP32 blocks
  // is about 8 u32 * 1.2 * rho_s ||A||_F for a loop that turns to FP32 storage
  // at relative residual rho_s.  The loop runs its first blocks in FP64 until
  // the squared relative residual is at most rstar2 = rho_s^2 with that noise
  // at 15% of tol (kNoiseFrac); with tol = 0, or a tol large enough that
  // rho_s >= 1, it is FP32 from the start.  The noise is not aligned with the
  // residual (measured: the stored sum exceeds the ID's true squared residual
  // by the noise energy, never falls short of it), so it inflates the tracked
  // sum by about kNoiseFrac^2 tol^2 = 0.0225 tol^2 near the stop.  The tracked
  // sum decides the stopping test except within (1 - kBandLo, 1 + kBandHi) tol^2,
  // where the FP64 recompute of the ID residual decides: the band holds that
  // inflation with room to spare, and a shortfall of up to 12% (a noise
  // aligned against the residual).
  constexpr double kNoiseFrac = 0.15, kNoisePerRho = 8.0 * 1.2, kU32 = 5.9604644775390625e-08;
  constexpr double kBandLo = 0.12, kBandHi = 0.1236;
  const double rstar = opts->tol > 0.0 ? kNoiseFrac * opts->tol / (kNoisePerRho * kU32) : HUGE_VAL;
  const bool phase64 = mixed && rstar < 1.0;  // FP64 storage first, FP32 from rstar2 on
  const double rstar2 = rstar * rstar;
  bool m32 = mixed && !phase64;               // the current block's storage is FP32
  // The FP64 loop on FP64 input reads A itself in its first block (no working
  // copy is made up front: the norm pass only reads A, the first block's
  // sampled columns are gathered from A, and its projection and update read
  // the trailing columns from A through the permutation, the update writing
  // the working copy).  This needs the 128-column update kernel (b <= 64).
  const bool direct = !m32 && !a_f32 && rbrp_gemm_trail_first_ok(mp, b, mp);
  RbrpWorkspace ws_static;
  {
    // (before the first launch: a workspace that cannot be had leaves every output untouched)
    const int64_t bb = std::max<int64_t>(b, 1);
    const bool w64 = !m32, vd = mixed || direct;
    const int e = noalloc ? static_workspace(h, mp, n, kmax, bb, mixed, w64, vd, &ws_static)
                          : ensure_workspace(h, mp, n, kmax, bb, mixed, w64, vd);
    if (e) return e;
  }
  RbrpWorkspace& ws = noalloc ? ws_static : h->ws;
  HIPCHK(rbrp_launch_iota(st, res->perm, n));
  int* perm = res->perm;
  const double tol2 = opts->tol * opts->tol;
  const uint64_t key0 = opts->seed, key1 = 0;
  const uint64_t ctr[4] = {opts->stream_offset, 0, 0, 0};
  const double ea = a_f32 ? 4.0 : 8.0;      // bytes per element of the input
  PhaseTimer pt(h, !noalloc);

  // R and the two block-state records (laid out right after it) start at zero
  HIPCHK(rbrp_launch_zero(st, ws.R, (int64_t)(((char*)(ws.st + 2) - (char*)ws.R) / sizeof(double))));
  int ph = pt.begin(RBRP_PHASE_NORMS);
  if (direct)
    HIPCHK(rbrp_launch_colnorm(st, (const double*)dA, lda, m, n, ws.diags));
  else if (m32)
    HIPCHK(rbrp_launch_copy_colnorm32(st, dA, a_f32 ? 1 : 0, lda, m, mp, n, ws.Aw32, ws.diags));
  else if (a_f32)
    HIPCHK(rbrp_launch_copy_colnorm_from32(st, (const float*)dA, lda, m, mp, n, ws.Aw, ws.diags));
  else
    HIPCHK(rbrp_launch_copy_colnorm(st, (const double*)dA, lda, m, mp, n, ws.Aw, ws.diags));
  pt.end(ph, 1, 2.0 * m * n, ea * m * n + (direct ? 0.0 : (m32 ? 4.0 : 8.0) * mp * n) + 8.0 * n);
  HIPCHK(rbrp_launch_sum_to_state(st, ws.diags, n, ws.st, 0));
  HIPCHK(hipMemcpyAsync(h->host_state, ws.st, sizeof(RbrpBlockState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // the device copy's seq is never written: mark the host copy as holding the
  // last publication handed out, so only the next one can match
  __atomic_store_n(&h->host_state->seq, h->pub_seq, __ATOMIC_RELEASE);
  pt.flush();
  const double N0 = h->host_state->N0;
  if (!std::isfinite(N0)) return RBRP_ERR_NUMERIC;

  // FP64 operands of the sampled block (gathered into Vd in the mixed loop
  // and in the direct first block, else read in place) and of the skeleton basis
  auto in_vd = [&](int64_t ib) { return m32 || (direct && ib == 0); };
  auto vblk = [&](int64_t ib) { return in_vd(ib) ? mat(ws.Vd, mp) : mat_cols(ws.Aw, mp, ws.J); };
  auto vkept = [&](int64_t ib) {
    return in_vd(ib) ? mat_cols(ws.Vd, mp, ws.Jkpos) : mat_cols(ws.Aw, mp, ws.Jk);
  };
  auto qall = [&]() { return m32 ? mat(ws.Q64, mp) : mat(ws.Aw, mp); };


  const int* d_s = &ws.st->s;
  const int* d_bt = &ws.st->bt;
  // the caller's matrix in the permuted order from column ib + bt on (direct first block)
  auto first_src = [&](int64_t ib) {
    return FirstSrc{(const double*)dA, lda, perm, ib, d_bt, m};
  };
  // Front of a block: sampling and the second Gram-Schmidt pass.  With `dyn`
  // the block's first column is ibase + *d_bt (the state still holds the
  // previous block's kept count): the front is queued before the previous
  // block's state reaches the host, sized for the largest first column
  // ibase + b, and does nothing (the sampler sets s = 0) if that block ends
  // the loop.  *ph_cgs receives the CGS phase record.
  auto queue_front = [&](int64_t ibase, bool dyn, int* ph_cgs) -> int {
    int ph = pt.begin(RBRP_PHASE_SAMPLE);
    HIPCHK(rbrp_launch_sample(st, ws.diags, ibase, n, kmax, b, key0, key1, ctr, ws.w, ws.cdf, ws.st,
                              ws.J, dyn ? d_bt : nullptr, tol2));
    pt.end(ph, 1, 0.0, 16.0 * (n - ibase));
    *ph_cgs = pt.begin(RBRP_PHASE_CGS);
    int cgs_launches = 0;
    if (m32) {
      HIPCHK(rbrp_launch_gather_v64(st, mp, b, ws.st, ws.Aw32, mp, ws.J, ws.Vd, mp));
      ++cgs_launches;
    } else if (direct && ibase == 0 && !dyn) {
      HIPCHK(rbrp_launch_gather_first(st, (const double*)dA, lda, m, mp, b, ws.st, ws.J, ws.Vd, mp));
      ++cgs_launches;
    }
    const int64_t imax = dyn ? std::min<int64_t>(ibase + b, kmax) : ibase;
    if (imax > 0) {
      const DimSpec di = dyn ? dim_dyn(ibase, d_bt, 1) : dim_static(ibase);
      int ns;
      int64_t kc;
      // C = Q^T V computed as its transpose V^T Q (the sampled block on the
      // 64-row side of the tiles), reduced into C column-major and added into
      // the sampled columns of R (R(0:i, J) += C)
      choose_split(b, imax, mp, b, ws.Pcap, &ns, &kc, 16, 128, h->tn64_resident);
      HIPCHK(rbrp_launch_gemm_tn64(st, dim_dyn(0, d_s, 1), di, b, imax, mp, ns, kc, vblk(imax), qall(), ws.P,
                                   b, imax * (int64_t)b));
      HIPCHK(rbrp_launch_reduce_splits(st, dim_dyn(0, d_s, 1), di, b, imax, ws.P, b,
                                       imax * (int64_t)b, ns, MatRef{nullptr, 0, nullptr, 0, nullptr},
                                       GEMM_SET, ws.Cbuf, ldr, ws.R, ldr, ws.J));
      HIPCHK(rbrp_launch_gemm_nn64(st, dim_static(mp), mp, dim_dyn(0, d_s, 1), b, di, qall(),
                                   mat(ws.Cbuf, ldr), vblk(imax), GEMM_SUB, h->nn64_resident));
      cgs_launches += 3;
    }
    pt.end(*ph_cgs, cgs_launches);
    return 0;
  };

  // Mixed loop: the FP64 squared ID residual ||A(:, perm[k:n]) - A(:, perm[0:k]) X||_F^2
  // with X = R11^{-1} R12 (the error the ID returns), for the stopping test.  In
  // column chunks: X's chunk by the blocked back substitution into the split-K
  // scratch, then E = A(:, rd) - A(:, J) X by the norms form of the tall NN
  // kernel reading A in place (FP64, or FP32 widened), which stores only
  // per-row-range squared column sums; all sums in a fixed order.  Synchronises
  // the stream; *out < 0 when a chunk's buffers cannot hold one column.
  auto recompute64 = [&](int64_t k, double* out) -> int {
    *out = -1.0;
    const int64_t nrem = n - k;
    if (nrem <= 0) {
      *out = 0.0;
      return 0;
    }
    if (k <= 0 || k > ws.Pcap) return 0;
    const int64_t npcap = ((mp + 63) / 64) * n;  // norm-partial capacity
    int64_t nc = std::min<int64_t>(nrem, ws.Pcap / k);
    while (nc > 1 && rbrp_gemm_nn64_ranges(m, nc, h->nn64_resident) * nc > npcap) nc = (nc + 1) / 2;
    if (rbrp_gemm_nn64_ranges(m, nc, h->nn64_resident) * nc > npcap) return 0;
    const int ph64 = pt.begin(RBRP_PHASE_RESID);
    int launches = 0;
    int64_t chunks = 0;
    double* const Aj = (double*)dA;  // (an FP32 matrix's float pointer when a_f32)
    for (int64_t c0 = 0; c0 < nrem; c0 += nc, ++chunks) {
      const int64_t ncc = std::min(nc, nrem - c0);
      HIPCHK(rbrp_launch_copy_block(st, k, ncc, ws.R + (k + c0) * ldr, ldr, ws.P, k));
      HIPCHK(rbrp_trsm_upper_left(st, k, ncc, ws.R, ldr, ws.P, k, &launches, ws.Dinv));
      HIPCHK(rbrp_launch_gemm_nn64_resnorm(st, m, ncc, k, MatRef{Aj, lda, perm, 0, nullptr},
                                           mat(ws.P, k), MatRef{Aj, lda, perm + k + c0, 0, nullptr},
                                           a_f32 ? 1 : 0, ws.normpart, ncc, h->nn64_resident));
      HIPCHK(rbrp_launch_sum(st, ws.normpart, rbrp_gemm_nn64_ranges(m, ncc, h->nn64_resident) * ncc,
                             ws.rcol + chunks));
      launches += 3;
    }
    HIPCHK(rbrp_launch_sum(st, ws.rcol, chunks, ws.r64));
    pt.end(ph64, launches + 1, 2.0 * m * k * nrem + (double)k * k * nrem, ea * m * n);
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, ws.r64, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *out = v;
    return 0;
  };

  int64_t i = 0, nb = 0;
  double err = (N0 > 0.0) ? 1.0 : 0.0;
  int ph_cgs = -1;
  if (i < kmax && N0 > 0.0) {
    const int e = queue_front(0, false, &ph_cgs);
    if (e != 0) return e;
  }
  while (i < kmax && N0 > 0.0) {
    // (this block's sampling and CGS are already on the stream)
    // ---- panel: Gram, pivoted Cholesky + filter, CholQR2 ----------------
    ph = pt.begin(RBRP_PHASE_PANEL);
    {
      // Gram products: the symmetric split-K kernel, one 64- or 128-column
      // tile per split.  A 128-column block is split over every resident
      // workgroup slot (two per compute unit: 1.7 ms less per ID at 10^6
      // rows than one per unit); narrower blocks as the other products.
      int ns;
      int64_t kc;
      if (b > 64) {
        const int64_t slots = h->tn64_resident > 0 ? h->tn64_resident : 512;
        int64_t n0 = std::min<int64_t>(slots, std::max<int64_t>(1, ws.Pcap / ((int64_t)b * b)));
        n0 = std::min<int64_t>(n0, std::max<int64_t>(1, mp / 64));
        kc = round_up((mp + n0 - 1) / n0, 16);
        ns = (int)((mp + kc - 1) / kc);
      } else {
        choose_split(b, b, mp, b, ws.Pcap, &ns, &kc);
      }
      auto gram = [&](DimSpec d, MatRef X) {
        return rbrp_launch_gram(st, d, b, mp, ns, kc, X, ws.P, b, (int64_t)b * b);
      };
      HIPCHK(gram(dim_dyn(0, d_s, 1), vblk(i)));
      HIPCHK(rbrp_launch_reduce_splits(st, dim_dyn(0, d_s, 1), dim_dyn(0, d_s, 1), b, b, ws.P, b,
                                       (int64_t)b * b, ns, mat(ws.G, b), GEMM_SET));
      HIPCHK(rbrp_launch_pivchol_filter(st, ws.G, b, b, opts->filter, i, ws.J, ws.st, ws.Rb, ws.Rinv,
                                        ws.Jk, ws.Jin, ws.Jout, ws.Jkpos));
      if (m32)  // the filtered-out columns' corrected values replace their stored residual
        HIPCHK(rbrp_launch_scatter_v32(st, mp, b, ws.st, ws.Vd, mp, ws.J, ws.Aw32, mp, ws.Jkpos));
      // Q1 = V(:, Jk) Rb^{-1}
      HIPCHK(rbrp_launch_gemm_nn64(st, dim_static(mp), mp, dim_dyn(0, d_bt, 1), b,
                                   dim_dyn(0, d_bt, 1), vkept(i), mat(ws.Rinv, b), mat(ws.Qbuf, mp),
                                   GEMM_SET, h->nn64_resident));
      HIPCHK(gram(dim_dyn(0, d_bt, 1), mat(ws.Qbuf, mp)));
      HIPCHK(rbrp_launch_reduce_splits(st, dim_dyn(0, d_bt, 1), dim_dyn(0, d_bt, 1), b, b, ws.P, b,
                                       (int64_t)b * b, ns, mat(ws.G2, b), GEMM_SET));
      // R2, R2^{-1} and R(i:i+bt, i:i+bt) = R2 Rb
      HIPCHK(rbrp_launch_chol2(st, ws.G2, b, b, ws.st, ws.Rb, ws.R2, ws.R2inv, ws.R + i + i * ldr,
                               ldr));
    }
    pt.end(ph, (b <= 64 ? 7 : 8) + (m32 ? 1 : 0));

    // ---- symmetric-difference permutation -------------------------------
    ph = pt.begin(RBRP_PHASE_PERMUTE);
    const int perm_launches = 2;  // the moves and the Q_b product
    if (m32)
      HIPCHK(rbrp_launch_perm_block32(st, i, b, ws.st, ws.Jk, ws.Jin, ws.Jout, perm, ws.diags, ws.R,
                                      ldr, ws.Aw32, mp, mp));
    else  // (the direct first block reads A through perm: no working-matrix move)
      HIPCHK(rbrp_launch_perm_block(st, i, b, ws.st, ws.Jk, ws.Jin, ws.Jout, perm, ws.diags, ws.R,
                                    ldr, (direct && i == 0) ? nullptr : ws.Aw, mp, mp));
    // Q_b = Q1 R2^{-1} into the block's columns of the skeleton basis
    // (mixed loop: with its FP32 copy in the working matrix's block columns)
    HIPCHK(rbrp_launch_gemm_nn64(st, dim_static(mp), mp, dim_dyn(0, d_bt, 1), b, dim_dyn(0, d_bt, 1),
                                 mat(ws.Qbuf, mp), mat(ws.R2inv, b),
                                 m32 ? mat_off(ws.Q64, mp, i) : mat_off(ws.Aw, mp, i), GEMM_SET,
                                 h->nn64_resident, m32 ? ws.Aw32 + i * mp : nullptr, mp));
    pt.end(ph, perm_launches);

    // ---- trailing projection and update ----------------------------------
    const int64_t nrem_max = n - i - 1;
    int ph_proj = -1, ph_upd = -1;
    int64_t groups = 0;  // norm partials written by the update (one per row group)
    if (nrem_max > 0) {
      int ns;
      int64_t kc;
      // (the FP32 kernel takes a block wider than 64 in one 128-row tile)
      const bool wide32 = m32 && b > 64;
      choose_split(b, nrem_max, mp, b, ws.Pcap, &ns, &kc, m32 ? 32 : 16, 128,
                   m32 ? (wide32 ? h->tn32w_resident : h->tn32_resident) : h->tn64_resident,
                   wide32 ? 128 : 64);
      hipEvent_t ev0, ev1;
      ph_proj = pt.begin_kernel(RBRP_PHASE_PROJECT, &ev0, &ev1);
      if (m32)
        HIPCHK(rbrp_launch_gemm_tn32(st, dim_dyn(0, d_bt, 1), dim_dyn(n - i, d_bt, -1), b, nrem_max,
                                     mp, ns, kc, ws.Aw32, i, nullptr, ws.Aw32, i, d_bt, mp, ws.P, b,
                                     (int64_t)b * nrem_max, ev0, ev1));
      else {
        const FirstSrc fs = first_src(i + 0);
        HIPCHK(rbrp_launch_gemm_tn64(st, dim_dyn(0, d_bt, 1), dim_dyn(n - i, d_bt, -1), b, nrem_max,
                                     mp, ns, kc, mat_off(ws.Aw, mp, i), mat_off(ws.Aw, mp, i, d_bt),
                                     ws.P, b, (int64_t)b * nrem_max, ev0, ev1,
                                     (direct && i == 0) ? &fs : nullptr));
      }
      pt.end(ph_proj, 1);
      ph = pt.begin(RBRP_PHASE_PROJECT_RED);
      if (m32)
        HIPCHK(rbrp_launch_reduce_splits_f32copy(st, dim_dyn(0, d_bt, 1), dim_dyn(n - i, d_bt, -1), b,
                                                 nrem_max, ws.P, b, (int64_t)b * nrem_max, ns,
                                                 mat_off(ws.R + i, ldr, i, d_bt), GEMM_SET, ws.Wt32,
                                                 nrem_max));
      else
        HIPCHK(rbrp_launch_reduce_splits(st, dim_dyn(0, d_bt, 1), dim_dyn(n - i, d_bt, -1), b,
                                         nrem_max, ws.P, b, (int64_t)b * nrem_max, ns,
                                         mat_off(ws.R + i, ldr, i, d_bt), GEMM_SET, ws.Wt, nrem_max));
      pt.end(ph, 1);
      ph_upd = pt.begin_kernel(RBRP_PHASE_UPDATE, &ev0, &ev1);
      if (m32)
        HIPCHK(rbrp_launch_gemm_trail32(st, mp, dim_dyn(n - i, d_bt, -1), nrem_max, dim_dyn(0, d_bt, 1),
                                        b, ws.Aw32, mp, i, nullptr, ws.Wt32, nrem_max, ws.Aw32, mp, i,
                                        d_bt, ws.normpart, nrem_max, &groups, ev0, ev1));
      else {
        const FirstSrc fs = first_src(i + 0);
        HIPCHK(rbrp_launch_gemm_trail(st, mp, dim_dyn(n - i, d_bt, -1), nrem_max, dim_dyn(0, d_bt, 1),
                                      b, ws.Aw, mp, i, nullptr, ws.Wt, nrem_max, ws.Aw, mp, i, d_bt,
                                      ws.normpart, nrem_max, &groups, ev0, ev1,
                                      (direct && i == 0) ? &fs : nullptr));
      }
      pt.end(ph_upd, 1);
    }
    // column norms, residual total and the state's publication to the host
    h->pub_seq = h->pub_seq == INT_MAX ? 1 : h->pub_seq + 1;
    const int seq = h->pub_seq;
    ph = pt.begin(RBRP_PHASE_RESID);
    if (nrem_max > 0)
      HIPCHK(rbrp_launch_colnorm_reduce(st, ws.normpart, nrem_max, groups, i, n, ws.st, ws.diags));
    HIPCHK(rbrp_launch_resid_total(st, ws.diags, i, n, ws.st, h->host_state, seq, ws.st_snap));
    pt.end(ph, nrem_max > 0 ? 2 : 1);

    // the next block's front goes on the stream before this block's state is
    // read back, so the device does not idle while the host decides
    const size_t next_mark = pt.recs.size();
    int ph_cgs_next = -1;
    {
      const int e = queue_front(i, true, &ph_cgs_next);
      if (e != 0) return e;
    }
    HIPCHK(wait_published(h, st, ws.st_snap, seq));
    const RbrpBlockState hs = *h->host_state;
    if (hs.status != 0) {
      pt.flush();
      return RBRP_ERR_NUMERIC;
    }
    if (hs.s <= 0 || hs.bt <= 0) {
      pt.recs.resize(next_mark);  // the front that does nothing
      break;
    }
    const int64_t bt = hs.bt, s = hs.s;
    const int64_t nrem = n - i - bt;
    if (h->timing) {
      // algorithmic work, now that the block's sizes are known
      const double ew = m32 ? 4.0 : 8.0;  // bytes per element of the working matrix
      pt.set_work(ph_cgs, i > 0 ? 4.0 * m * i * s : 0.0, i > 0 ? 2.0 * 8.0 * m * i + 16.0 * m * s : 0.0);
      pt.set_work(ph_proj, 2.0 * m * bt * nrem, ew * (m * nrem + m * bt) + 8.0 * bt * nrem);
      pt.set_work(ph_upd, 2.0 * m * bt * nrem, ew * (2.0 * m * nrem + m * bt + bt * nrem));
    }
    // the records stay until the ID ends (no event waits between blocks)
    res->draws = hs.draws;
    err = hs.err;
    i += bt;
    bool stop = err <= tol2 || i >= kmax;
    // FP32 storage: the sum of the stored residual carries the rounding noise
    // of the FP32 updates (at most kNoiseFrac tol by the schedule above), so
    // where that noise could decide the test the FP64 recompute of the ID
    // residual decides, and is the error recorded for the block.
    bool requeue = false;
    const double band_lo = (1.0 - kBandLo) * tol2, band_hi = (1.0 + kBandHi) * tol2;
    if (m32 && tol2 > 0.0 && i < kmax && i < n && err <= band_hi && err > band_lo) {
      double r64;
      const int e = recompute64(i, &r64);
      if (e != 0) return e;
      if (r64 >= 0.0) {
        err = r64 / N0;
        // the queued front did nothing when the FP32 sum was already below tol^2
        requeue = err > tol2 && hs.err <= tol2;
        stop = err <= tol2;
      }
    }
    if (nb < res->trace_cap) {
      res->ranks[nb] = i;
      res->error[nb] = err;
      res->t_ranks[nb] = now_s() - t_start;
    }
    ++nb;
    if (stop) {
      pt.recs.resize(next_mark);  // the front that does nothing or is not needed
      break;
    }
    if (requeue) {
      pt.recs.resize(next_mark);
      const int e = queue_front(i, false, &ph_cgs_next);
      if (e != 0) return e;
    }
    if (phase64 && !m32 && err <= rstar2) {
      // FP32 storage from the next block on (its front, already queued, has
      // re-orthogonalised its sampled columns in the FP64 working matrix):
      // the FP64 skeleton basis, the FP32 residual with the norms of the stored
      // values, and the sampled block's FP64 copy for the panel
      ph = pt.begin(RBRP_PHASE_NORMS);
      HIPCHK(rbrp_launch_copy_block(st, mp, i, ws.Aw, mp, ws.Q64, mp));
      HIPCHK(rbrp_launch_copy_colnorm32(st, ws.Aw + i * mp, 0, mp, mp, mp, n - i, ws.Aw32 + i * mp,
                                        ws.diags + i));
      HIPCHK(rbrp_launch_gather_first(st, ws.Aw, mp, mp, mp, b, ws.st, ws.J, ws.Vd, mp));
      pt.end(ph, 3, 0.0, 8.0 * mp * n + 4.0 * mp * (n - i));
      m32 = true;
    }
    ph_cgs = ph_cgs_next;
  }
  const int64_t k = i;
  res->k = k;
  res->nblocks = nb;
  res->err_final = std::sqrt(std::max(err, 0.0));

  // ---- X = R11^{-1} R12 (matlab/driver.m:86) --------------------------------
  if (res->X && k > 0 && n > k) {
    ph = pt.begin(RBRP_PHASE_SOLVE);
    const int64_t nc = n - k;
    HIPCHK(rbrp_launch_copy_block(st, k, nc, ws.R + k * ldr, ldr, res->X, res->ldx));
    int launches = 1;
    HIPCHK(rbrp_trsm_upper_left(st, k, nc, ws.R, ldr, res->X, res->ldx, &launches, ws.Dinv));
    pt.end(ph, launches, (double)k * k * nc, 8.0 * (0.5 * k * k + 2.0 * k * nc));
  }
  HIPCHK(hipStreamSynchronize(st));
  pt.flush();
  res->t_total = now_s() - t_start;
  return 0;
}

}  // namespace

int rbrp_colid_impl(rbrp_handle_t h, const void* dA, bool a_f32, int64_t m, int64_t n, int64_t lda,
                    const rbrp_opts_t* opts, rbrp_result_t* res) {
  return colid_impl(h, dA, a_f32, m, n, lda, opts, res);
}

extern "C" {

int rbrp_colid_d(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda,
                 const rbrp_opts_t* opts, rbrp_result_t* res) {
  return colid_impl(h, dA, false, m, n, lda, opts, res);
}

int rbrp_colid_s(rbrp_handle_t h, const float* dA, int64_t m, int64_t n, int64_t lda,
                 const rbrp_opts_t* opts, rbrp_result_t* res) {
  return colid_impl(h, dA, true, m, n, lda, opts, res);
}

int rbrp_colid_d_noalloc(rbrp_handle_t h, const double* dA, int64_t m, int64_t n, int64_t lda,
                         const rbrp_opts_t* opts, rbrp_result_t* res) {
  return colid_impl(h, dA, false, m, n, lda, opts, res, true);
}

}  // extern "C"
