// Host side of the batched column ID (rbrp_colid_batched_d, rbrp_batched_tune).
//
// Persistent mode: the problem descriptors and a largest-first queue order are
// uploaded into device metadata, every workgroup of one launch owns a
// workspace slot (working matrix, R, CholQR2 and permutation buffers sized for
// the largest problem) and pulls problems from the queue counter until it is
// drained.  The launch (queue reset + persistent kernel) is recorded once in a
// hipGraph on a private stream and replayed by later calls with the same launch
// shape: only the metadata upload and the result download surround it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/rbrp.h"
#include "rbrp_batched.h"
#include "rbrp_handle.h"

int rbrp_colid_impl(rbrp_handle_t h, const void* dA, bool a_f32, int64_t m, int64_t n, int64_t lda,
                    const rbrp_opts_t* opts, rbrp_result_t* res);

static_assert(sizeof(rbrp_problem_t) == sizeof(RbrpBatchProblem), "problem layout");
static_assert(sizeof(rbrp_problem_result_t) == sizeof(RbrpBatchResult), "result layout");

#define HIPCHK(x)                         \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) return (int)e_; \
  } while (0)

namespace {

inline size_t up256(size_t x) { return (x + 255) / 256 * 256; }

struct BatchPlan {
  RbrpBatchShape shape;
  int64_t m_max = 0, n_max = 0;
  int fits = 1;  // persistent kernel limits hold
};

int validate(int64_t nprob, const rbrp_problem_t* pr, const rbrp_opts_t* o, BatchPlan* plan) {
  if (nprob < 0) return -2;
  if (nprob > 0 && !pr) return -3;
  if (!o) return -4;
  if (o->b < 1 || o->b > 128) return -4;
  if (!(o->tol >= 0.0) || o->tol > 1e300) return -4;
  if (o->max_rank < 0) return -4;
  if (o->precision != RBRP_PREC_FP64) return -4;
  if (o->filter != RBRP_FILTER_SAMPLED && o->filter != RBRP_FILTER_NOMINAL) return -4;
  int64_t kmax_max = 1;
  for (int64_t q = 0; q < nprob; ++q) {
    const rbrp_problem_t& p = pr[q];
    if (p.m < 0 || p.n < 0 || p.m > 0x7fffffffLL || p.n > 0x7fffffffLL) return -3;
    if (p.lda < std::max<int64_t>(p.m, 1)) return -3;
    if (p.m > 0 && p.n > 0 && !p.A) return -3;
    if (p.n > 0 && !p.perm) return -3;
    const int64_t pmn = std::min(p.m, p.n);
    const int64_t kmax = o->max_rank > 0 ? std::min<int64_t>(o->max_rank, pmn) : pmn;
    if (p.X && p.ldx < std::max<int64_t>(kmax, 1)) return -3;
    plan->m_max = std::max(plan->m_max, p.m);
    plan->n_max = std::max(plan->n_max, p.n);
    kmax_max = std::max(kmax_max, kmax);
  }
  plan->fits = (plan->n_max <= RBRP_BATCH_MAXN && o->b <= RBRP_BATCH_MAXB) ? 1 : 0;
  RbrpBatchShape& sh = plan->shape;
  const int64_t mp = (plan->m_max + 15) / 16 * 16;
  const int64_t nm = std::max<int64_t>(plan->n_max, 1);
  sh.ldr = kmax_max;
  size_t off = up256(sizeof(double) * (size_t)std::max<int64_t>(mp, 16) * nm);
  sh.off_R = off;
  off += up256(sizeof(double) * (size_t)sh.ldr * nm);
  sh.off_Q = off;
  off += up256(sizeof(double) * (size_t)std::max<int64_t>(mp, 16) * o->b);
  sh.off_C = off;
  off += up256(sizeof(double) * (size_t)sh.ldr * o->b);
  sh.off_T = off;
  off += up256(sizeof(double) * (size_t)sh.ldr * 2 * o->b);
  sh.slot_bytes = off;
  sh.n_max = (int)nm;
  sh.b = (int)o->b;
  sh.filter = o->filter;
  sh.max_rank = (int)std::min<int64_t>(o->max_rank, 0x7fffffff);
  sh.tol2 = o->tol * o->tol;
  sh.ws = nullptr;
  sh.tile = 2;
  return 0;
}

int grow(void** p, size_t* have, size_t need, bool host) {
  if (*have >= need) return 0;
  if (*p) {
    if (host)
      (void)hipHostFree(*p);
    else
      (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
  }
  hipError_t e = host ? hipHostMalloc(p, need, hipHostMallocDefault) : hipMalloc(p, need);
  if (e != hipSuccess) {
    *p = nullptr;
    return RBRP_ERR_ALLOC;
  }
  *have = need;
  return 0;
}

void drop_graph(rbrp_handle_t h) {
  if (h->bt_exec) (void)hipGraphExecDestroy(h->bt_exec);
  if (h->bt_graph) (void)hipGraphDestroy(h->bt_graph);
  h->bt_exec = nullptr;
  h->bt_graph = nullptr;
  h->bt_key.clear();
}

std::vector<uint64_t> tune_key(const BatchPlan& pl) {
  return {(uint64_t)pl.m_max, (uint64_t)pl.n_max, (uint64_t)pl.shape.b};
}

int compute_units(rbrp_handle_t h, int* cus, size_t* lds_max) {
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, h->device));
  *cus = std::max(1, prop.multiProcessorCount);
  *lds_max = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                   : (size_t)64 * 1024;
  return 0;
}

// Uploads descriptors + queue order, runs the persistent launch (through the
// recorded graph when use_graph), downloads the results.  ms (optional)
// receives the device time of the launch itself.
int run_persistent(rbrp_handle_t h, int64_t nprob, const rbrp_problem_t* pr, BatchPlan& pl,
                   int wg_per_cu, int tile, bool use_graph, rbrp_problem_result_t* results,
                   float* ms, int* per_cu_used = nullptr) {
  int cus;
  size_t lds_cu;
  {
    int e = compute_units(h, &cus, &lds_cu);
    if (e) return e;
  }
  (void)lds_cu;
  HIPCHK(rbrp_batched_prepare(pl.shape.n_max, pl.shape.b));
  pl.shape.tile = tile;
  // persistent workgroups beyond what is co-resident would only start once
  // the queue is drained: cap at the occupancy limit
  const int fit = rbrp_batched_max_blocks_per_cu(pl.shape.n_max, pl.shape.b, tile);
  const int per_cu = std::max(1, std::min(wg_per_cu, fit));
  if (per_cu_used) *per_cu_used = per_cu;
  const int grid = (int)std::min<int64_t>((int64_t)per_cu * cus, nprob);
  // workspace and metadata
  {
    int e = grow(&h->bt_ws, &h->bt_ws_bytes, pl.shape.slot_bytes * (size_t)grid, false);
    if (e) return e;
  }
  const size_t sz_p = up256(sizeof(RbrpBatchProblem) * (size_t)nprob);
  const size_t sz_o = up256(sizeof(int) * (size_t)nprob);
  const size_t sz_q = 256;
  const size_t sz_r = up256(sizeof(RbrpBatchResult) * (size_t)nprob);
  const size_t meta = sz_p + sz_o + sz_q + sz_r;
  bool realloc = h->bt_meta_bytes < meta;
  {
    int e = grow(&h->bt_meta, &h->bt_meta_bytes, meta, false);
    if (e) return e;
    e = grow(&h->bt_host, &h->bt_host_bytes, sz_p + sz_o + sz_r, true);
    if (e) return e;
  }
  if (realloc) drop_graph(h);
  pl.shape.ws = h->bt_ws;
  char* dm = (char*)h->bt_meta;
  RbrpBatchProblem* d_probs = (RbrpBatchProblem*)dm;
  int* d_order = (int*)(dm + sz_p);
  int* d_queue = (int*)(dm + sz_p + sz_o);
  RbrpBatchResult* d_res = (RbrpBatchResult*)(dm + sz_p + sz_o + sz_q);
  char* hm = (char*)h->bt_host;
  std::memcpy(hm, pr, sizeof(RbrpBatchProblem) * (size_t)nprob);
  int* h_order = (int*)(hm + sz_p);
  std::iota(h_order, h_order + nprob, 0);
  std::stable_sort(h_order, h_order + nprob, [&](int a, int c) {
    return pr[a].m * pr[a].n > pr[c].m * pr[c].n;  // largest first
  });
  if (!h->bt_stream) HIPCHK(hipStreamCreateWithFlags(&h->bt_stream, hipStreamNonBlocking));
  hipStream_t st = h->bt_stream;
  HIPCHK(hipStreamSynchronize(h->stream));  // order after earlier work of the handle
  HIPCHK(hipMemcpyAsync(d_probs, hm, sizeof(RbrpBatchProblem) * (size_t)nprob,
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_order, h_order, sizeof(int) * (size_t)nprob, hipMemcpyHostToDevice, st));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
  }
  if (use_graph) {
    const RbrpBatchShape& s = pl.shape;
    std::vector<uint64_t> key = {(uint64_t)grid,       (uint64_t)nprob,       (uint64_t)(uintptr_t)s.ws,
                                 (uint64_t)(uintptr_t)dm, s.slot_bytes,        s.off_R,
                                 s.off_Q,              s.off_C,               s.off_T,
                                 (uint64_t)s.ldr,      (uint64_t)s.n_max,     (uint64_t)s.b,
                                 (uint64_t)s.filter,   (uint64_t)s.max_rank,  (uint64_t)s.tile,
                                 0};
    std::memcpy(&key.back(), &s.tol2, sizeof(double));
    if (!h->bt_exec || h->bt_key != key) {
      drop_graph(h);
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      hipError_t le = rbrp_launch_colid_batched(st, grid, d_probs, d_order, (int)nprob, d_queue, d_res,
                                                pl.shape);
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(st, &g);
      if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return (int)le;
      }
      HIPCHK(ce);
      h->bt_graph = g;
      HIPCHK(hipGraphInstantiate(&h->bt_exec, g, nullptr, nullptr, 0));
      h->bt_key = key;
    }
    HIPCHK(hipGraphLaunch(h->bt_exec, st));
  } else {
    HIPCHK(rbrp_launch_colid_batched(st, grid, d_probs, d_order, (int)nprob, d_queue, d_res,
                                     pl.shape));
  }
  if (ms) HIPCHK(hipEventRecord(e1, st));
  RbrpBatchResult* h_res = (RbrpBatchResult*)(hm + sz_p + sz_o);
  HIPCHK(hipMemcpyAsync(h_res, d_res, sizeof(RbrpBatchResult) * (size_t)nprob,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (ms) {
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  std::memcpy(results, h_res, sizeof(RbrpBatchResult) * (size_t)nprob);
  return 0;
}

int run_loop(rbrp_handle_t h, int64_t nprob, const rbrp_problem_t* pr, const rbrp_opts_t* o,
             rbrp_problem_result_t* results) {
  for (int64_t q = 0; q < nprob; ++q) {
    const rbrp_problem_t& p = pr[q];
    rbrp_opts_t oq = *o;
    oq.seed = p.seed;
    oq.stream_offset = p.stream_offset;
    rbrp_result_t r{};
    r.perm = p.perm;
    r.X = p.X;
    r.ldx = p.ldx;
    r.trace_cap = 0;
    const int st = rbrp_colid_impl(h, p.A, false, p.m, p.n, p.lda, &oq, &r);
    rbrp_problem_result_t& out = results[q];
    std::memset(&out, 0, sizeof(out));
    if (st == RBRP_ERR_NUMERIC) {
      out.status = 1;
      continue;
    }
    if (st) return st;
    out.k = r.k;
    out.nblocks = r.nblocks;
    out.err_final = r.err_final;
    out.draws = r.draws;
  }
  return 0;
}

// The launch shape for this batch shape: the tuner's pick on this handle, or
// four workgroups per CU and two row tiles.  Packed as wg_per_cu | tile << 8.
int launch_shape(rbrp_handle_t h, const BatchPlan& pl) {
  const std::vector<uint64_t> key = tune_key(pl);
  for (const auto& kv : h->bt_tuned)
    if (kv.first == key) return kv.second;
  return 4 | (2 << 8);
}

}  // namespace

void rbrp_batched_release(rbrp_handle_t h) {
  if (!h) return;
  if (h->bt_stream) (void)hipStreamSynchronize(h->bt_stream);
  drop_graph(h);
  if (h->bt_ws) (void)hipFree(h->bt_ws);
  if (h->bt_meta) (void)hipFree(h->bt_meta);
  if (h->bt_host) (void)hipHostFree(h->bt_host);
  if (h->bt_stream) (void)hipStreamDestroy(h->bt_stream);
  h->bt_ws = h->bt_meta = h->bt_host = nullptr;
  h->bt_ws_bytes = h->bt_meta_bytes = h->bt_host_bytes = 0;
  h->bt_stream = nullptr;
}

extern "C" {

int rbrp_colid_batched_d(rbrp_handle_t h, int64_t nprob, const rbrp_problem_t* problems,
                         const rbrp_opts_t* opts, rbrp_problem_result_t* results, int mode) {
  if (!h) return -1;
  if (mode < RBRP_BATCH_AUTO || mode > RBRP_BATCH_LOOP) return -6;
  BatchPlan pl;
  {
    int e = validate(nprob, problems, opts, &pl);
    if (e) return e;
  }
  if (nprob > 0 && !results) return -5;
  if (nprob == 0) return 0;
  if (mode == RBRP_BATCH_PERSISTENT && !pl.fits) return -4;
  if (mode == RBRP_BATCH_LOOP || !pl.fits) return run_loop(h, nprob, problems, opts, results);
  const int shp = launch_shape(h, pl);
  return run_persistent(h, nprob, problems, pl, shp & 0xff, shp >> 8, true, results, nullptr);
}

int rbrp_batched_tune(rbrp_handle_t h, int64_t nprob, const rbrp_problem_t* problems,
                      const rbrp_opts_t* opts, int* wg_per_cu) {
  if (!h) return -1;
  BatchPlan pl;
  {
    int e = validate(nprob, problems, opts, &pl);
    if (e) return e;
  }
  if (!wg_per_cu) return -5;
  if (nprob == 0 || !pl.fits) return -4;
  std::vector<rbrp_problem_result_t> res((size_t)nprob);
  // launch shapes: row tiles per wave of the fused projection + update (the
  // register footprint changes with it, and with it the occupancy) x
  // workgroups per compute unit
  const int cand[] = {1, 2, 3, 4, 6, 8};
  float best = 0.f;
  int pick = 0, pick_tile = 2;
  for (int tile : {1, 2, 4}) {
    int last_used = 0;
    for (int c : cand) {
      float ms = 0.f;
      BatchPlan p2 = pl;
      int used = 0;
      int e = run_persistent(h, nprob, problems, p2, c, tile, false, res.data(), &ms, &used);
      if (e) return e;
      if (used == last_used) continue;  // capped to the same launch as the previous candidate
      last_used = used;
      if (pick == 0 || ms < best) {
        best = ms;
        pick = used;
        pick_tile = tile;
      }
    }
  }
  const std::vector<uint64_t> key = tune_key(pl);
  const int shp = pick | (pick_tile << 8);
  bool found = false;
  for (auto& kv : h->bt_tuned)
    if (kv.first == key) {
      kv.second = shp;
      found = true;
    }
  if (!found) h->bt_tuned.push_back({key, shp});
  *wg_per_cu = pick;
  return 0;
}

int rbrp_batched_launch_shape(rbrp_handle_t h, int64_t nprob, const rbrp_problem_t* problems,
                              const rbrp_opts_t* opts, int* wg_per_cu, int* tile_rows) {
  if (!h) return -1;
  BatchPlan pl;
  {
    int e = validate(nprob, problems, opts, &pl);
    if (e) return e;
  }
  if (!wg_per_cu || !tile_rows) return -5;
  const int shp = launch_shape(h, pl);
  *wg_per_cu = shp & 0xff;
  *tile_rows = shp >> 8;
  return 0;
}

}  // extern "C"
