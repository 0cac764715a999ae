// batch.hip -- many small graphs counted by one device pipeline.
//
// Layout in HBM.  The graphs of a batch are stored twice:
//   * union form: one block-diagonal CSR over all graphs (graph g owns vertex
//     ids [vbase_g, vbase_g + nv_g) and entries [ebase_g, ebase_g + ne_g)),
//     plus gid[v] = graph of vertex v.  The ordinary orientation kernels and
//     the persistent intersection kernel run over it once, whatever the number
//     of graphs: a tile (2048 oriented edges) maps to a (graph, edge range)
//     piece of the batch, tiles are drawn from a device work queue, short-list
//     edges are intersected by threads and long-list edges are binned to the
//     wave-per-edge kernel.  Per-graph totals are reduced in LDS per tile.
//   * local form: each graph's own CSR (ids 0..nv_g-1), used by the baseline
//     that launches the single-graph pipeline once per graph.
// All scratch is sized at creation so the pipeline can be captured once as a
// hipGraph and replayed (no allocation, no memset node inside the capture).
// Graphs added from host memory are staged: xt_batch_add copies the local CSR
// into a pinned host mirror of the local arena and returns; the next launch
// (or device-side generation) flushes all staged graphs with two bulk
// asynchronous copies and one rebase kernel over all of them.
#include <cstring>
#include <vector>

#include "xt_device.hpp"
#include "xt_internal.hpp"

namespace xt {

int gen_er_rows(xt_ctx* ctx, Scratch& s, int64_t n, uint64_t seed, uint64_t thr64, int64_t* off,
                int32_t* tails, int64_t cap, int64_t* ne_out, hipStream_t st);

__global__ __launch_bounds__(256) void k_rebase(int64_t nv, int64_t ne, int64_t vbase, int64_t ebase,
                                                int32_t g, const int64_t* __restrict__ loff,
                                                const int32_t* __restrict__ ltails,
                                                int64_t* __restrict__ goff, int32_t* __restrict__ gtails,
                                                int32_t* __restrict__ gid) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t n = nv + 1 > ne ? nv + 1 : ne;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (i <= nv) goff[vbase + i] = ebase + loff[i];
    if (i < nv) gid[vbase + i] = g;
    if (i < ne) gtails[ebase + i] = (int32_t)(vbase + ltails[i]);
  }
}

// Rebase of many staged graphs in one launch.  rec holds kRecWords words per
// graph {nv, ne, vbase, ebase, lobase, work start}; graph j spans work items
// [start_j, start_j + max(nv_j + 1, ne_j)).  Each thread takes a run of
// kRebaseRun consecutive items: one search, then a forward walk.
constexpr int kRecWords = 6;
constexpr int kRebaseRun = 8;

__global__ __launch_bounds__(256) void k_rebase_many(int64_t ng, int64_t total, int32_t g0,
                                                     const int64_t* __restrict__ rec,
                                                     const int64_t* __restrict__ loff,
                                                     const int32_t* __restrict__ ltails,
                                                     int64_t* __restrict__ goff,
                                                     int32_t* __restrict__ gtails,
                                                     int32_t* __restrict__ gid) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kRebaseRun;
  for (int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kRebaseRun; i0 < total;
       i0 += stride) {
    int64_t lo = 0, hi = ng - 1;  // last graph with start <= i0
    while (lo < hi) {
      const int64_t mid = lo + ((hi - lo + 1) >> 1);
      if (rec[mid * kRecWords + 5] <= i0)
        lo = mid;
      else
        hi = mid - 1;
    }
    int64_t j = lo;
    const int64_t i1 = i0 + kRebaseRun < total ? i0 + kRebaseRun : total;
    for (int64_t i = i0; i < i1; i++) {
      while (j + 1 < ng && rec[(j + 1) * kRecWords + 5] <= i) j++;
      const int64_t* r = rec + j * kRecWords;
      const int64_t nv = r[0], ne = r[1], vbase = r[2], ebase = r[3], lobase = r[4];
      const int64_t k = i - r[5];
      if (k <= nv) goff[vbase + k] = ebase + loff[lobase + k];
      if (k < nv) gid[vbase + k] = g0 + (int32_t)j;
      if (k < ne) gtails[ebase + k] = (int32_t)(vbase + ltails[ebase + k]);
    }
  }
}

__global__ void k_copy_total(const uint64_t* __restrict__ counters, uint64_t* __restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = counters[kTotal];
}

}  // namespace xt

using namespace xt;

struct BatchGraph {
  int64_t nv, ne, vbase, ebase, lobase;
};

struct xt_batch {
  xt_ctx* ctx = nullptr;
  int64_t max_v = 0, max_e = 0, max_g = 0;
  int64_t V = 0, E = 0, LO = 0;
  std::vector<BatchGraph> graphs;
  int64_t* goff = nullptr;
  int32_t* gtails = nullptr;
  int32_t* gid = nullptr;
  int64_t* loff = nullptr;
  int32_t* ltails = nullptr;
  uint64_t* gtotal = nullptr;
  uint64_t* ptotal = nullptr;
  Scratch s;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t exec_V = -1, exec_E = -1, exec_G = -1;
  uint32_t exec_flags = 0;
  int exec_small_cut = -1;
  int exec_group_blocks = -1;
  int exec_group_round = -1;
  int exec_items_per_tile = -1;
  int exec_tile_run = -1;
  int exec_wave_owner = -2;
  int exec_group_hash = 0;
  // pinned host mirror of the local arena for staged host adds
  int64_t* h_loff = nullptr;
  int32_t* h_ltails = nullptr;
  int64_t* h_rec = nullptr;  // kRecWords per staged graph
  uint64_t* h_totals = nullptr;  // pinned per-graph totals copied back by each launch (batch_readback_reserve)
  int64_t* d_rec = nullptr;
  int64_t pend_g = 0;         // graphs [pend_g, graphs.size()) are staged, not yet on the device
  hipEvent_t ev_stage = nullptr;
  bool stage_inflight = false;
  int last_mode = -1;
  bool running = false;
  xt_batch_timing timing = {};
};

namespace xt {

static void batch_free(xt_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void* ps[] = {b->goff, b->gtails, b->gid, b->loff, b->ltails, b->gtotal, b->ptotal, b->d_rec};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  void* hs[] = {b->h_loff, b->h_ltails, b->h_rec, b->h_totals};
  for (void* p : hs)
    if (p) (void)hipHostFree(p);
  if (b->ev_stage) (void)hipEventDestroy(b->ev_stage);
  release(b->s);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

static int batch_check_room(xt_batch* b, int64_t nv, int64_t ne) {
  if ((int64_t)b->graphs.size() >= b->max_g || b->V + nv > b->max_v || b->E + ne > b->max_e ||
      b->LO + nv + 1 > b->max_v + b->max_g)
    return set_err(b->ctx, XT_ERR_CAPACITY, "batch capacity exceeded");
  return XT_OK;
}

// Appends the local CSR already placed at loff+LO / ltails+E and builds its union form.
static int batch_commit(xt_batch* b, int64_t nv, int64_t ne) {
  const int32_t g = (int32_t)b->graphs.size();
  BatchGraph rec{nv, ne, b->V, b->E, b->LO};
  const int64_t n = (nv + 1 > ne ? nv + 1 : ne);
  int64_t grid = (n + 255) / 256;
  if (grid > 4096) grid = 4096;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_rebase, dim3((unsigned)grid), dim3(256), 0, b->stream, nv, ne, rec.vbase,
                     rec.ebase, g, (const int64_t*)(b->loff + rec.lobase),
                     (const int32_t*)(b->ltails + rec.ebase), b->goff, b->gtails, b->gid);
  XT_HIP(b->ctx, hipGetLastError());
  b->graphs.push_back(rec);
  b->V += nv;
  b->E += ne;
  b->LO += nv + 1;
  b->pend_g = (int64_t)b->graphs.size();
  return XT_OK;
}

// Waits until the last flush has finished reading the pinned staging buffers.
static int batch_stage_idle(xt_batch* b) {
  if (b->stage_inflight) {
    XT_HIP(b->ctx, hipEventSynchronize(b->ev_stage));
    b->stage_inflight = false;
  }
  return XT_OK;
}

// Moves every staged graph to the device: two bulk copies of the local CSR
// ranges (the staged graphs form a suffix of the local arena) and one rebase
// launch for all of them.  Stream-ordered; nothing waits on the host.
static int batch_flush(xt_batch* b) {
  const int64_t G = (int64_t)b->graphs.size();
  if (b->pend_g >= G) return XT_OK;
  xt_ctx* ctx = b->ctx;
  XT_TRY(batch_stage_idle(b));
  const BatchGraph& first = b->graphs[b->pend_g];
  const int64_t lo0 = first.lobase, e0 = first.ebase;
  XT_HIP(ctx, hipMemcpyAsync(b->loff + lo0, b->h_loff + lo0, (size_t)(b->LO - lo0) * sizeof(int64_t),
                             hipMemcpyHostToDevice, b->stream));
  if (b->E > e0)
    XT_HIP(ctx, hipMemcpyAsync(b->ltails + e0, b->h_ltails + e0, (size_t)(b->E - e0) * sizeof(int32_t),
                               hipMemcpyHostToDevice, b->stream));
  const int64_t ng = G - b->pend_g;
  int64_t work = 0;
  for (int64_t j = 0; j < ng; j++) {
    const BatchGraph& r = b->graphs[b->pend_g + j];
    int64_t* w = b->h_rec + j * kRecWords;
    w[0] = r.nv;
    w[1] = r.ne;
    w[2] = r.vbase;
    w[3] = r.ebase;
    w[4] = r.lobase;
    w[5] = work;
    work += r.nv + 1 > r.ne ? r.nv + 1 : r.ne;
  }
  XT_HIP(ctx, hipMemcpyAsync(b->d_rec, b->h_rec, (size_t)ng * kRecWords * sizeof(int64_t),
                             hipMemcpyHostToDevice, b->stream));
  int64_t grid = (work + 256 * kRebaseRun - 1) / (256 * kRebaseRun);
  if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(k_rebase_many, dim3((unsigned)grid), dim3(256), 0, b->stream, ng, work,
                     (int32_t)b->pend_g, (const int64_t*)b->d_rec, (const int64_t*)b->loff,
                     (const int32_t*)b->ltails, b->goff, b->gtails, b->gid);
  XT_HIP(ctx, hipGetLastError());
  XT_HIP(ctx, hipEventRecord(b->ev_stage, b->stream));
  b->stage_inflight = true;
  b->pend_g = G;
  return XT_OK;
}

// Launches the union pipeline on the batch stream (no synchronisation).
static int batch_pipeline(xt_batch* b, uint32_t flags) {
  xt_ctx* ctx = b->ctx;
  hipStream_t st = b->stream;
  const int mode = (flags & XT_ORIENT_ID) ? kOrientId : kOrientDegree;
  const int64_t G = (int64_t)b->graphs.size();
  XT_TRY(fill_u64(ctx, b->s.counters.as<uint64_t>(), kNumSlots, 0, st));
  XT_TRY(fill_u64(ctx, b->gtotal, G, 0, st));
  XT_TRY(orient<int32_t>(ctx, b->s, b->V, b->E, b->goff, b->gtails, mode, st));
  XT_TRY(count_oriented<int32_t>(ctx, b->s, b->V, b->E, b->goff, b->gtails, nullptr, b->gid, b->gtotal,
                                 st, nullptr));
  return XT_OK;
}

int batch_launch(xt_batch* b, int mode, uint32_t flags) {
  xt_ctx* ctx = b->ctx;
  hipStream_t st = b->stream;
  const int64_t G = (int64_t)b->graphs.size();
  b->last_mode = mode;
  if (G == 0) return XT_OK;
  XT_TRY(batch_flush(b));
  if (mode == XT_BATCH_PERSISTENT) {
    XT_HIP(ctx, hipEventRecord(b->ev0, st));
    XT_TRY(batch_pipeline(b, flags));
    if (b->h_totals)  // the totals follow the kernels: done with the launch's last event
      XT_HIP(ctx, hipMemcpyAsync(b->h_totals, b->gtotal, (size_t)G * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipEventRecord(b->ev1, st));
  } else if (mode == XT_BATCH_HIPGRAPH) {
    if (!b->exec || b->exec_V != b->V || b->exec_E != b->E || b->exec_G != G ||
        b->exec_flags != flags || b->exec_small_cut != ctx->small_cut ||
        b->exec_group_blocks != ctx->group_blocks || b->exec_group_round != ctx->group_round ||
        b->exec_items_per_tile != ctx->items_per_tile || b->exec_tile_run != ctx->tile_run ||
        b->exec_wave_owner != ctx->wave_owner || b->exec_group_hash != ctx->group_hash) {
      if (b->exec) {
        XT_HIP(ctx, hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
      }
      if (b->graph) {
        XT_HIP(ctx, hipGraphDestroy(b->graph));
        b->graph = nullptr;
      }
      XT_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = batch_pipeline(b, flags);
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(st, &gr);
      if (rc != XT_OK) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      if (e != hipSuccess) return hip_err(ctx, e, "hipStreamEndCapture");
      b->graph = gr;
      XT_HIP(ctx, hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0));
      b->exec_V = b->V;
      b->exec_E = b->E;
      b->exec_G = G;
      b->exec_flags = flags;
      b->exec_small_cut = ctx->small_cut;
      b->exec_group_blocks = ctx->group_blocks;
      b->exec_group_round = ctx->group_round;
      b->exec_items_per_tile = ctx->items_per_tile;
      b->exec_tile_run = ctx->tile_run;
      b->exec_wave_owner = ctx->wave_owner;
      b->exec_group_hash = ctx->group_hash;
    }
    XT_HIP(ctx, hipEventRecord(b->ev0, st));
    XT_HIP(ctx, hipGraphLaunch(b->exec, st));
    XT_HIP(ctx, hipEventRecord(b->ev1, st));
  } else if (mode == XT_BATCH_PER_GRAPH) {
    XT_HIP(ctx, hipEventRecord(b->ev0, st));
    for (int64_t g = 0; g < G; g++) {
      const BatchGraph& r = b->graphs[g];
      XT_TRY(count_device_csr<int32_t>(ctx, b->s, r.nv, r.ne, b->loff + r.lobase, b->ltails + r.ebase,
                                       flags & XT_ORIENT_ID, st, false));
      hipLaunchKernelGGL(k_copy_total, dim3(1), dim3(64), 0, st,
                         (const uint64_t*)b->s.counters.as<uint64_t>(), b->ptotal + g);
      XT_HIP(ctx, hipGetLastError());
    }
    XT_HIP(ctx, hipEventRecord(b->ev1, st));
  } else {
    return set_err(ctx, XT_ERR_ARG, "unknown batch mode");
  }
  b->running = true;
  return XT_OK;
}

// 1 when the last launch has finished, 0 while it runs, <0 on error.
int batch_poll(xt_batch* b) {
  if (!b->running) return 1;
  hipError_t e = hipEventQuery(b->ev1);
  if (e == hipErrorNotReady) return 0;
  if (e != hipSuccess) return hip_err(b->ctx, e, "hipEventQuery");
  float ms = 0;
  XT_HIP(b->ctx, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->timing.run_ms = ms;
  b->running = false;
  return 1;
}

int batch_wait(xt_batch* b) {
  if (!b->running) return XT_OK;
  XT_HIP(b->ctx, hipEventSynchronize(b->ev1));
  return batch_poll(b) < 0 ? XT_ERR_HIP : XT_OK;
}

// Pinned memory that each persistent launch copies the per-graph totals into
// (after its kernels, on its stream), so a finished launch needs no further
// device call to read them (batch_read).
int batch_readback_reserve(xt_batch* b) {
  if (b->h_totals) return XT_OK;
  hipError_t e = hipHostMalloc((void**)&b->h_totals, (size_t)(b->max_g > 0 ? b->max_g : 1) * sizeof(uint64_t));
  if (e != hipSuccess) return hip_err(b->ctx, e, "batch readback buffer");
  return XT_OK;
}

int batch_read(xt_batch* b, uint64_t* totals, int64_t n) {
  const int64_t G = (int64_t)b->graphs.size();
  if (n > G) n = G;
  if (n <= 0) return XT_OK;
  if (b->h_totals && b->last_mode == XT_BATCH_PERSISTENT && !b->running) {  // copied back by the launch
    std::memcpy(totals, b->h_totals, (size_t)n * sizeof(uint64_t));
    return XT_OK;
  }
  const uint64_t* src = b->last_mode == XT_BATCH_PER_GRAPH ? b->ptotal : b->gtotal;
  XT_HIP(b->ctx, hipMemcpyAsync(totals, src, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
  XT_HIP(b->ctx, hipStreamSynchronize(b->stream));
  return XT_OK;
}

// Checks a host CSR graph of local ids and, when off_out/tails_out are given,
// copies it in the same pass; an error message, or nullptr.  The loops are
// branch-free so the compiler vectorises them.
const char* check_host_graph(int64_t nv, int64_t ne, const int64_t* off, const int32_t* tails,
                             int64_t* off_out, int32_t* tails_out) {
  if (nv < 1 || ne < 0 || !off || (ne > 0 && !tails)) return "bad graph";
  if (off[0] != 0 || off[nv] != ne) return "offsets must start at 0 and end at ne";
  int64_t down = 0;
  if (off_out) {
    off_out[0] = 0;
    for (int64_t i = 0; i < nv; i++) {
      const int64_t a = off[i], b = off[i + 1];
      off_out[i + 1] = b;
      down |= (int64_t)(b < a);
    }
  } else {
    for (int64_t i = 0; i < nv; i++) down |= (int64_t)(off[i + 1] < off[i]);
  }
  if (down) return "offsets are not monotone";
  uint32_t out = 0;
  const uint32_t n = (uint32_t)nv;  // nv < 2^31: negative ids wrap above it
  if (tails_out) {
    for (int64_t e = 0; e < ne; e++) {
      const int32_t v = tails[e];
      tails_out[e] = v;
      out |= (uint32_t)((uint32_t)v >= n);
    }
  } else {
    for (int64_t e = 0; e < ne; e++) out |= (uint32_t)((uint32_t)tails[e] >= n);
  }
  if (out) return "neighbour id out of range [0, nv)";
  return nullptr;
}

// Allocates the pinned staging mirror (once per batch).
int batch_stage_reserve(xt_batch* b) {
  if (b->h_loff) return XT_OK;
  hipError_t e = hipHostMalloc((void**)&b->h_loff, (size_t)(b->max_v + b->max_g) * sizeof(int64_t));
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_ltails, (size_t)b->max_e * sizeof(int32_t));
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_rec, (size_t)b->max_g * kRecWords * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_rec, (size_t)b->max_g * kRecWords * sizeof(int64_t));
  if (e == hipSuccess && !b->ev_stage) e = hipEventCreateWithFlags(&b->ev_stage, hipEventDisableTiming);
  if (e != hipSuccess) return hip_err(b->ctx, e, "batch staging buffers");
  return XT_OK;
}

// Stages a graph that check_host_graph accepted: copies it into the pinned
// mirror; the next launch moves it to the device.
int batch_stage_host(xt_batch* b, int64_t nv, int64_t ne, const int64_t* off, const int32_t* tails) {
  XT_TRY(batch_check_room(b, nv, ne));
  XT_TRY(batch_stage_reserve(b));
  XT_TRY(batch_stage_idle(b));
  std::memcpy(b->h_loff + b->LO, off, (size_t)(nv + 1) * sizeof(int64_t));
  if (ne > 0) std::memcpy(b->h_ltails + b->E, tails, (size_t)ne * sizeof(int32_t));
  b->graphs.push_back(BatchGraph{nv, ne, b->V, b->E, b->LO});
  b->V += nv;
  b->E += ne;
  b->LO += nv + 1;
  return XT_OK;
}

// Reserves one graph's place in the pinned mirror (offsets at *off_dst, ids at
// *tails_dst) without copying it; the caller fills both before the next launch.
int batch_stage_layout(xt_batch* b, int64_t nv, int64_t ne, int64_t** off_dst, int32_t** tails_dst) {
  XT_TRY(batch_check_room(b, nv, ne));
  XT_TRY(batch_stage_reserve(b));
  XT_TRY(batch_stage_idle(b));
  *off_dst = b->h_loff + b->LO;
  *tails_dst = b->h_ltails + b->E;
  b->graphs.push_back(BatchGraph{nv, ne, b->V, b->E, b->LO});
  b->V += nv;
  b->E += ne;
  b->LO += nv + 1;
  return XT_OK;
}

// Checks a host graph and copies it into the pinned mirror in the same pass
// (the serving path's borrowed requests).  Returns the check's message (the
// graph is then not staged) or nullptr; *rc carries a staging error.
const char* batch_stage_host_checked(xt_batch* b, int64_t nv, int64_t ne, const int64_t* off,
                                     const int32_t* tails, int* rc) {
  *rc = batch_check_room(b, nv, ne);
  if (*rc == XT_OK) *rc = batch_stage_reserve(b);
  if (*rc == XT_OK) *rc = batch_stage_idle(b);
  if (*rc != XT_OK) return nullptr;
  const char* bad = check_host_graph(nv, ne, off, tails, b->h_loff + b->LO, b->h_ltails + b->E);
  if (bad) return bad;
  b->graphs.push_back(BatchGraph{nv, ne, b->V, b->E, b->LO});
  b->V += nv;
  b->E += ne;
  b->LO += nv + 1;
  return nullptr;
}

int batch_add_host(xt_batch* b, int64_t nv, int64_t ne, const int64_t* off, const int32_t* tails) {
  if (nv >= (1LL << 31) - 1) return set_err(b->ctx, XT_ERR_ARG, "nv must be below 2^31-1");
  const char* bad = check_host_graph(nv, ne, off, tails, nullptr, nullptr);
  if (bad) return set_err(b->ctx, nv < 1 || ne < 0 || !off ? XT_ERR_ARG : XT_ERR_INPUT, bad);
  return batch_stage_host(b, nv, ne, off, tails);
}

}  // namespace xt

extern "C" {

int xt_batch_create(xt_ctx* ctx, int64_t max_vertices, int64_t max_edges, int64_t max_graphs,
                    xt_batch** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out || max_vertices < 1 || max_edges < 1 || max_graphs < 1) return XT_ERR_ARG;
    if (max_vertices >= (1LL << 31) - 1) return set_err(ctx, XT_ERR_ARG, "max_vertices must be below 2^31-1");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    xt_batch* b = new xt_batch();
    b->ctx = ctx;
    b->max_v = max_vertices;
    b->max_e = max_edges;
    b->max_g = max_graphs;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipMalloc(&b->goff, (size_t)(max_vertices + 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&b->gtails, (size_t)max_edges * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&b->gid, (size_t)max_vertices * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&b->loff, (size_t)(max_vertices + max_graphs) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&b->ltails, (size_t)max_edges * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&b->gtotal, (size_t)max_graphs * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&b->ptotal, (size_t)max_graphs * sizeof(uint64_t));
    if (e != hipSuccess) {
      int rc = hip_err(ctx, e, "xt_batch_create");
      batch_free(b);
      return rc;
    }
    int rc = reserve(ctx, b->s, max_vertices, max_edges, false);
    if (rc == XT_OK) rc = fill_u64(ctx, b->s.counters.as<uint64_t>(), kNumSlots, 0, b->stream);
    if (rc == XT_OK) {
      e = hipMemsetAsync(b->goff, 0, sizeof(int64_t), b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess) rc = hip_err(ctx, e, "xt_batch_create");
    }
    if (rc != XT_OK) {
      batch_free(b);
      return rc;
    }
    *out = b;
    return XT_OK;
  });
}

int xt_batch_add(xt_batch* b, int64_t nv, int64_t ne, const int64_t* offsets, const int32_t* tails,
                 int64_t* slot) {
  return xt_guarded([&]() -> int {
    if (!b) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    XT_HIP(b->ctx, hipSetDevice(b->ctx->device));
    if (b->running) return set_err(b->ctx, XT_ERR_STATE, "batch is running");
    XT_TRY(batch_add_host(b, nv, ne, offsets, tails));
    if (slot) *slot = (int64_t)b->graphs.size() - 1;
    return XT_OK;
  });
}

int xt_batch_add_generated(xt_batch* b, uint64_t base_seed, int64_t first, int64_t count, int64_t nmin,
                           int64_t nmax) {
  return xt_guarded([&]() -> int {
    if (!b || first < 0 || count < 0) return XT_ERR_ARG;
    xt_ctx* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (b->running) return set_err(ctx, XT_ERR_STATE, "batch is running");
    for (int64_t g = first; g < first + count; g++) {
      int64_t n;
      int kind;
      uint64_t gseed, thr;
      XT_TRY(xt_batch_params(base_seed, g, nmin, nmax, &n, &kind, &gseed, &thr));
      if (kind == 0) {
        // random geometric: host cell-grid generator (reference LCG stream per graph)
        std::vector<int64_t> off(n + 1);
        int64_t m = 0;
        const uint32_t s32 = (uint32_t)(gseed ^ (gseed >> 32));
        XT_TRY(xt_rgg_generate(n, s32, 0.0, off.data(), nullptr, 0, &m));
        std::vector<int64_t> t64(m > 0 ? m : 1);
        XT_TRY(xt_rgg_generate(n, s32, 0.0, off.data(), t64.data(), m, &m));
        std::vector<int32_t> t32(t64.begin(), t64.begin() + m);
        XT_TRY(batch_add_host(b, n, m, off.data(), t32.data()));
      } else {
        // Erdos-Renyi: generated on the device straight into the local arena,
        // behind any staged host graphs
        XT_TRY(batch_flush(b));
        XT_TRY(batch_check_room(b, n, 0));
        int64_t m = 0;
        XT_TRY(gen_er_rows(ctx, b->s, n, gseed + 3, thr, b->loff + b->LO, nullptr, 0, &m, b->stream));
        XT_TRY(batch_check_room(b, n, m));
        XT_TRY(gen_er_rows(ctx, b->s, n, gseed + 3, thr, b->loff + b->LO, b->ltails + b->E, m, &m,
                           b->stream));
        XT_TRY(batch_commit(b, n, m));
      }
    }
    XT_TRY(batch_flush(b));
    XT_HIP(ctx, hipStreamSynchronize(b->stream));
    return XT_OK;
  });
}

int xt_batch_run(xt_batch* b, int mode, uint32_t flags) {
  return xt_guarded([&]() -> int {
    if (!b) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    XT_HIP(b->ctx, hipSetDevice(b->ctx->device));
    if (flags & XT_PER_VERTEX) return set_err(b->ctx, XT_ERR_ARG, "batches report per-graph totals only");
    XT_TRY(batch_launch(b, mode, flags));
    XT_TRY(batch_wait(b));
    b->timing.n_graphs = (int64_t)b->graphs.size();
    b->timing.n_vertices = b->V;
    b->timing.n_edges = b->E;
    b->timing.n_tiles = (b->E / 2 + kChunk - 1) / kChunk;
    return XT_OK;
  });
}

int xt_batch_results(xt_batch* b, uint64_t* totals, int64_t n) {
  return xt_guarded([&]() -> int {
    if (!b || (n > 0 && !totals)) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    XT_HIP(b->ctx, hipSetDevice(b->ctx->device));
    XT_TRY(batch_wait(b));
    return batch_read(b, totals, n);
  });
}

int xt_batch_stats_get(xt_batch* b, uint32_t flags, xt_graph_stats* out) {
  return xt_guarded([&]() -> int {
    if (!b || !out) return XT_ERR_ARG;
    xt_ctx* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    XT_TRY(batch_wait(b));
    XT_TRY(batch_flush(b));
    *out = xt_graph_stats{};
    const int64_t G = (int64_t)b->graphs.size();
    if (G == 0) return XT_OK;
    hipStream_t st = b->stream;
    const int mode = (flags & XT_ORIENT_ID) ? kOrientId : kOrientDegree;
    // The union's orientation is each graph's own (graphs share no edge).
    XT_TRY(fill_u64(ctx, b->s.counters.as<uint64_t>(), kNumSlots, 0, st));
    XT_TRY(orient<int32_t>(ctx, b->s, b->V, b->E, b->goff, b->gtails, mode, st));
    XT_TRY(alg_stats(ctx, b->s, b->V, b->E, st,
                     ctx->small_cut > 0 ? ctx->small_cut
                     : b->E >= kBatchWaveMinEntries ? kBatchLightCut : kBatchLightCutSmall));
    uint64_t h[kNumSlots];
    int64_t nep = 0;
    XT_HIP(ctx, hipMemcpyAsync(h, b->s.counters.p, sizeof(h), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipMemcpyAsync(&nep, b->s.ooff.as<int64_t>() + b->V, sizeof(nep), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
    xt_graph_stats r = {};
    r.sum_min = h[kStatMin];
    r.ne_plus = nep;
    // per graph 8(n+1) + 4|E+| + 4 sum_min, summed over the graphs
    r.alg_bytes = 8ULL * (uint64_t)(b->V + G) + 4ULL * (uint64_t)nep + 4ULL * r.sum_min;
    r.sum_min_heavy = h[kStatMinHeavy];
    r.heavy_edges = (int64_t)h[kStatHeavyEdges];
    r.group_bytes = 4ULL * (r.sum_min_heavy + (uint64_t)r.heavy_edges);
    *out = r;
    return XT_OK;
  });
}

int xt_batch_timing_get(const xt_batch* b, xt_batch_timing* out) {
  return xt_guarded([&]() -> int {
    if (!b || !out) return XT_ERR_ARG;
    *out = b->timing;
    return XT_OK;
  });
}

int xt_batch_clear(xt_batch* b) {
  return xt_guarded([&]() -> int {
    if (!b) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    XT_TRY(batch_wait(b));
    XT_TRY(batch_stage_idle(b));
    b->graphs.clear();
    b->V = b->E = b->LO = 0;
    b->pend_g = 0;
    return XT_OK;
  });
}

int xt_batch_destroy(xt_batch* b) {
  return xt_guarded([&]() -> int {
    if (!b) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    batch_free(b);
    return XT_OK;
  });
}

}  // extern "C"
