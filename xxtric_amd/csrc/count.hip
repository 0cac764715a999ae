// count.hip -- intersection and reduction kernels over the oriented CSR (gfx950).
//
// Every oriented edge u->v closes one triangle per common out-neighbour w of u
// and v.  The shorter of the two out-lists is walked and each of its ids is
// searched in the longer one (lower_bound restarted from the previous hit, so
// a walk over a sorted list never searches backwards).
//
//  k_count_small: one thread per oriented edge (chunks of 2048 edges per
//                 256-thread workgroup, strided so a wave reads consecutive
//                 edges).  Edges whose shorter list exceeds `small_cut` are
//                 appended to a device list instead.
//  k_count_wave:  one wave64 per listed edge, lanes striding the shorter list.
//  Reduction:     per-thread uint64 -> wave64 shuffle tree -> LDS -> one
//                 64-bit atomic per workgroup.
//
// Per-vertex credit (XT_PER_VERTEX): the triangle {u, v, w} belongs to
// min(u, v, w) (xxtric.hpp:165).  Hits with w > min(u, v) are summed in a
// register and credited with one atomic per edge; hits with w < min(u, v)
// are credited to w directly.
#include "xt_device.hpp"
#include "xt_internal.hpp"

namespace xt {

struct EdgeLists {
  int64_t s0, s1, l0, l1;  // shorter [s0,s1), longer [l0,l1) in otails
};

__device__ __forceinline__ EdgeLists edge_lists(const int64_t* __restrict__ ooff, int64_t a0,
                                                int64_t a1, int64_t v) {
  const int64_t b0 = ooff[v], b1 = ooff[v + 1];
  EdgeLists L;
  if (a1 - a0 <= b1 - b0) {
    L.s0 = a0; L.s1 = a1; L.l0 = b0; L.l1 = b1;
  } else {
    L.s0 = b0; L.s1 = b1; L.l0 = a0; L.l1 = a1;
  }
  return L;
}

// Per-graph totals of a batch: a tile touches few graphs, so hits are summed
// in LDS slots relative to the tile's first graph and flushed once per tile.
constexpr int kGraphSlots = 64;

template <bool PV, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_count_small(
    int64_t nv, const int64_t* __restrict__ ooff, const int32_t* __restrict__ ot, int small_cut,
    int64_t defer_cap, uint64_t* __restrict__ counters, int64_t* __restrict__ defer,
    unsigned long long* __restrict__ pv, const int32_t* __restrict__ gid,
    unsigned long long* __restrict__ gtotal) {
  __shared__ int64_t s_off[kOffCap + 1];
  __shared__ int64_t s_lohi[2];
  __shared__ uint64_t s_red[kBlock / kWave];
  __shared__ unsigned long long s_g[BATCH ? kGraphSlots : 1];
  __shared__ int64_t s_tile;
  const int64_t nep = ooff[nv];
  const int64_t ntiles = (nep + kChunk - 1) / kChunk;
  for (;;) {
    // Tiles come from blockIdx (one tile per workgroup) or, for batches, from
    // a device work queue drained by a fixed grid of persistent workgroups.
    int64_t tile;
    if constexpr (BATCH) {
      if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd((unsigned long long*)&counters[kTileQueue], 1ull);
      if (threadIdx.x < kGraphSlots) s_g[threadIdx.x] = 0;
      __syncthreads();
      tile = s_tile;
    } else {
      tile = blockIdx.x;
    }
    if (tile >= ntiles) break;  // uniform across the workgroup
    const int64_t eb = tile * kChunk;
    const int64_t ee = min(nep, eb + (int64_t)kChunk);
    if (threadIdx.x == 0) {
      const int64_t lo = row_of(ooff, 0, nv - 1, eb);
      s_lohi[0] = lo;
      s_lohi[1] = row_of(ooff, lo, nv - 1, ee - 1);
    }
    __syncthreads();
    const int64_t ulo = s_lohi[0], uhi = s_lohi[1], span = uhi - ulo + 1;
    const bool in_lds = span <= kOffCap;
    if (in_lds)
      for (int64_t k = threadIdx.x; k <= span; k += kBlock) s_off[k] = ooff[ulo + k];
    const int32_t g0 = BATCH ? gid[ulo] : 0;
    __syncthreads();

    uint64_t hits_all = 0;
#pragma unroll 1
    for (int j = 0; j < kPerThread; j++) {
      const int64_t e = eb + (int64_t)j * kBlock + threadIdx.x;
      if (e >= ee) break;
      const int64_t u = in_lds ? ulo + lds_row_of(s_off, (int)span, e) : row_of(ooff, ulo, uhi, e);
      const int64_t a0 = in_lds ? s_off[u - ulo] : ooff[u];
      const int64_t a1 = in_lds ? s_off[u - ulo + 1] : ooff[u + 1];
      const int64_t v = ot[e];
      const EdgeLists L = edge_lists(ooff, a0, a1, v);
      const int64_t ns = L.s1 - L.s0;
      if (ns == 0) continue;
      if (ns > small_cut) {
        const unsigned long long idx = atomicAdd((unsigned long long*)&counters[kDeferred], 1ull);
        if ((int64_t)idx < defer_cap) {
          defer[2 * idx] = e;
          defer[2 * idx + 1] = u;
          continue;
        }
        // list full: fall through and count the edge here
      }
      const int64_t mmin = u < v ? u : v;
      uint64_t hits = 0, at_min = 0;
      int64_t pos = L.l0;
      const int32_t lmax = ot[L.l1 - 1];
      for (int64_t k = L.s0; k < L.s1; k++) {
        const int32_t w = ot[k];
        if (w > lmax) break;
        pos = lower_bound_dev<int32_t>(ot, pos, L.l1, w);
        if (ot[pos] == w) {
          hits++;
          if constexpr (PV) {
            if ((int64_t)w < mmin)
              atomicAdd(&pv[w], 1ull);
            else
              at_min++;
          }
          pos++;
          if (pos >= L.l1) break;
        }
      }
      if constexpr (PV) {
        if (at_min) atomicAdd(&pv[mmin], (unsigned long long)at_min);
      }
      if constexpr (BATCH) {
        if (hits) {
          const int32_t slot = gid[u] - g0;
          if (slot < kGraphSlots)
            atomicAdd(&s_g[slot], (unsigned long long)hits);
          else
            atomicAdd(&gtotal[gid[u]], (unsigned long long)hits);
        }
      } else {
        hits_all += hits;
      }
    }
    if constexpr (BATCH) {
      __syncthreads();
      if (threadIdx.x < kGraphSlots && s_g[threadIdx.x])
        atomicAdd(&gtotal[g0 + threadIdx.x], s_g[threadIdx.x]);
      __syncthreads();  // s_g and s_tile are reused by the next tile
    } else {
      const uint64_t bsum = block_sum_u64<kBlock>(hits_all, s_red);
      if (threadIdx.x == 0 && bsum)
        atomicAdd((unsigned long long*)&counters[kTotal], (unsigned long long)bsum);
      break;
    }
  }
}

template <bool PV, bool BATCH>
__global__ __launch_bounds__(256) void k_count_wave(const int64_t* __restrict__ ooff,
                                                    const int32_t* __restrict__ ot,
                                                    int64_t defer_cap,
                                                    uint64_t* __restrict__ counters,
                                                    const int64_t* __restrict__ defer,
                                                    unsigned long long* __restrict__ pv,
                                                    const int32_t* __restrict__ gid,
                                                    unsigned long long* __restrict__ gtotal) {
  __shared__ uint64_t s_red[256 / kWave];
  int64_t nd = (int64_t)counters[kDeferred];
  if (nd > defer_cap) nd = defer_cap;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wid = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * (256 / kWave);
  uint64_t hits_all = 0;
  for (int64_t d = wid; d < nd; d += nw) {  // uniform per wave
    const int64_t e = defer[2 * d];
    const int64_t u = defer[2 * d + 1];
    const int64_t v = ot[e];
    const EdgeLists L = edge_lists(ooff, ooff[u], ooff[u + 1], v);
    const int64_t mmin = u < v ? u : v;
    uint64_t hits = 0, at_min = 0;
    for (int64_t k = L.s0 + lane; k < L.s1; k += kWave) {
      const int32_t w = ot[k];
      const int64_t pos = lower_bound_dev<int32_t>(ot, L.l0, L.l1, w);
      if (pos < L.l1 && ot[pos] == w) {
        hits++;
        if constexpr (PV) {
          if ((int64_t)w < mmin)
            atomicAdd(&pv[w], 1ull);
          else
            at_min++;
        }
      }
    }
    if constexpr (PV) {
      at_min = wave_sum_u64(at_min);
      if (lane == 0 && at_min) atomicAdd(&pv[mmin], (unsigned long long)at_min);
    }
    if constexpr (BATCH) {
      hits = wave_sum_u64(hits);
      if (lane == 0 && hits) atomicAdd(&gtotal[gid[u]], (unsigned long long)hits);
    } else {
      hits_all += hits;
    }
  }
  if constexpr (!BATCH) {
    const uint64_t bsum = block_sum_u64<256>(hits_all, s_red);
    if (threadIdx.x == 0 && bsum) atomicAdd((unsigned long long*)&counters[kTotal], (unsigned long long)bsum);
  }
}

// Sum over oriented edges of min(d+(u), d+(v)) -- the algorithmic probe count.
__global__ __launch_bounds__(kBlock) void k_alg_stats(int64_t nv, const int64_t* __restrict__ ooff,
                                                      const int32_t* __restrict__ ot,
                                                      uint64_t* __restrict__ counters) {
  __shared__ int64_t s_lohi[2];
  __shared__ uint64_t s_red[kBlock / kWave];
  const int64_t nep = ooff[nv];
  const int64_t eb = (int64_t)blockIdx.x * kChunk;
  if (eb >= nep) return;
  const int64_t ee = min(nep, eb + (int64_t)kChunk);
  if (threadIdx.x == 0) {
    const int64_t lo = row_of(ooff, 0, nv - 1, eb);
    s_lohi[0] = lo;
    s_lohi[1] = row_of(ooff, lo, nv - 1, ee - 1);
  }
  __syncthreads();
  uint64_t acc = 0;
  for (int j = 0; j < kPerThread; j++) {
    const int64_t e = eb + (int64_t)j * kBlock + threadIdx.x;
    if (e >= ee) break;
    const int64_t u = row_of(ooff, s_lohi[0], s_lohi[1], e);
    const int64_t v = ot[e];
    const int64_t du = ooff[u + 1] - ooff[u], dv = ooff[v + 1] - ooff[v];
    acc += (uint64_t)(du < dv ? du : dv);
  }
  const uint64_t bsum = block_sum_u64<kBlock>(acc, s_red);
  if (threadIdx.x == 0 && bsum) atomicAdd((unsigned long long*)&counters[kStatMin], (unsigned long long)bsum);
}

static int64_t defer_capacity(int64_t ne_bound) {
  return ne_bound < (int64_t)(1 << 26) ? (ne_bound > 0 ? ne_bound : 1) : (int64_t)(1 << 26);
}

int count_oriented(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, uint64_t* pv,
                   const int32_t* gid, uint64_t* gtotal, hipStream_t st) {
  if (nv <= 0 || ne_bound <= 0) return XT_OK;
  const bool batch = gid != nullptr;
  const int64_t nchunks = (ne_bound + kChunk - 1) / kChunk;
  // Deferred-edge list: 2 int64 per edge, capped; overflow is counted inline.
  const int64_t defer_cap = defer_capacity(ne_bound);
  XT_TRY(ensure(ctx, s.defer, (size_t)defer_cap * 2 * sizeof(int64_t)));
  uint64_t* counters = s.counters.as<uint64_t>();
  XT_TRY(fill_u64(ctx, counters + kDeferred, 2, 0, st));
  if (batch) XT_TRY(fill_u64(ctx, counters + kTileQueue, 1, 0, st));
  const int64_t* ooff = s.ooff.as<int64_t>();
  const int32_t* ot = s.otails.as<int32_t>();
  int64_t* defer = s.defer.as<int64_t>();
  unsigned long long* pvp = (unsigned long long*)pv;
  unsigned long long* gt = (unsigned long long*)gtotal;
  const unsigned wave_grid = (unsigned)(ctx->num_cus * 8);
  int64_t persist = (int64_t)ctx->num_cus * ctx->persist_blocks;
  const unsigned small_grid = (unsigned)(batch ? (nchunks < persist ? nchunks : persist) : nchunks);
#define XT_COUNT_LAUNCH(P, B)                                                                         \
  do {                                                                                                \
    hipLaunchKernelGGL((k_count_small<P, B>), dim3(small_grid), dim3(kBlock), 0, st, nv, ooff, ot,    \
                       ctx->small_cut, defer_cap, counters, defer, pvp, gid, gt);                     \
    XT_HIP(ctx, hipGetLastError());                                                                   \
    hipLaunchKernelGGL((k_count_wave<P, B>), dim3(wave_grid), dim3(256), 0, st, ooff, ot, defer_cap, \
                       counters, (const int64_t*)defer, pvp, gid, gt);                                \
    XT_HIP(ctx, hipGetLastError());                                                                   \
  } while (0)
  if (pv && batch)
    XT_COUNT_LAUNCH(true, true);
  else if (pv)
    XT_COUNT_LAUNCH(true, false);
  else if (batch)
    XT_COUNT_LAUNCH(false, true);
  else
    XT_COUNT_LAUNCH(false, false);
#undef XT_COUNT_LAUNCH
  return XT_OK;
}

int alg_stats(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, hipStream_t st) {
  if (nv <= 0 || ne_bound <= 0) return XT_OK;
  const int64_t nchunks = (ne_bound + kChunk - 1) / kChunk;
  uint64_t* counters = s.counters.as<uint64_t>();
  XT_TRY(fill_u64(ctx, counters + kStatMin, 1, 0, st));
  hipLaunchKernelGGL(k_alg_stats, dim3((unsigned)nchunks), dim3(kBlock), 0, st, nv,
                     (const int64_t*)s.ooff.as<int64_t>(), (const int32_t*)s.otails.as<int32_t>(),
                     counters);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

int reserve(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne, bool per_vertex) {
  const int64_t nvv = nv > 0 ? nv : 1, nee = ne > 0 ? ne : 1;
  const int64_t nchunks = (nee + kChunk - 1) / kChunk;
  XT_TRY(ensure(ctx, s.deg, (size_t)nvv * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.dplus, (size_t)nvv * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.tailcnt, (size_t)nchunks * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.qscan, (size_t)(nchunks + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.ooff, (size_t)(nvv + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.otails, (size_t)nee * sizeof(int32_t)));
  XT_TRY(ensure(ctx, s.counters, kNumSlots * sizeof(uint64_t)));
  XT_TRY(ensure(ctx, s.defer, (size_t)defer_capacity(nee) * 2 * sizeof(int64_t)));
  const size_t st = scan_tmp_bytes(nvv) > scan_tmp_bytes(nchunks) ? scan_tmp_bytes(nvv) : scan_tmp_bytes(nchunks);
  XT_TRY(ensure(ctx, s.scan_tmp, st));
  if (per_vertex) XT_TRY(ensure(ctx, s.pv, (size_t)nvv * sizeof(uint64_t)));
  return XT_OK;
}

void release(Scratch& s) {
  DevBuf* bufs[] = {&s.deg, &s.dplus, &s.tailcnt, &s.qscan, &s.ooff, &s.otails, &s.counters,
                    &s.defer, &s.pv, &s.scan_tmp, &s.in_off, &s.in_tails};
  for (DevBuf* b : bufs) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->bytes = 0;
  }
}
// ---------------------------------------------------------------------------
// Reference lower-FFI shapes.

// counting.cu:477-533 semantics: pair i = (e1, m); v = L[m]; count n in (m, e1)
// with L[n] in N(v).  One wave64 per pair, lanes over n, binary search in N(v).
__global__ __launch_bounds__(256) void k_masked(int64_t nedges, const int64_t* __restrict__ mask,
                                                const int64_t* __restrict__ ind,
                                                const int64_t* __restrict__ L,
                                                unsigned long long* __restrict__ out) {
  __shared__ uint64_t s_red[256 / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wid = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * (256 / kWave);
  uint64_t acc = 0;
  for (int64_t i = wid; i < nedges; i += nw) {
    const int64_t e1 = mask[2 * i], m = mask[2 * i + 1];
    const int64_t v = L[m];
    const int64_t j0 = ind[v], j1 = ind[v + 1];
    if (j1 <= j0) continue;
    for (int64_t n = m + 1 + lane; n < e1; n += kWave) {
      const int64_t w = L[n];
      const int64_t p = lower_bound_dev<int64_t>(L, j0, j1, w);
      if (p < j1 && L[p] == w) acc++;
    }
  }
  const uint64_t b = block_sum_u64<256>(acc, s_red);
  if (threadIdx.x == 0 && b) atomicAdd(out, (unsigned long long)b);
}

// counting.cu:720-767 semantics: count i in [start, end) with L[i] in L[jstart, jend).
__global__ __launch_bounds__(256) void k_inner(int64_t start, int64_t end, int64_t jstart, int64_t jend,
                                               const int64_t* __restrict__ L,
                                               unsigned long long* __restrict__ out) {
  __shared__ uint64_t s_red[256 / kWave];
  uint64_t acc = 0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = start + (int64_t)blockIdx.x * 256 + threadIdx.x; i < end; i += stride) {
    const int64_t w = L[i];
    const int64_t p = lower_bound_dev<int64_t>(L, jstart, jend, w);
    if (p < jend && L[p] == w) acc++;
  }
  const uint64_t b = block_sum_u64<256>(acc, s_red);
  if (threadIdx.x == 0 && b) atomicAdd(out, (unsigned long long)b);
}

__global__ void k_add_total(const uint64_t* __restrict__ counters, unsigned long long* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicAdd(out, (unsigned long long)counters[kTotal]);
}

int launch_masked(xt_ctx* ctx, int64_t nedges, const int64_t* mask, const int64_t* ind,
                  const int64_t* L, int64_t* out, hipStream_t st) {
  if (nedges <= 0) return XT_OK;
  int64_t g = (nedges + 3) / 4;
  if (g > ctx->num_cus * 16) g = ctx->num_cus * 16;
  hipLaunchKernelGGL(k_masked, dim3((unsigned)g), dim3(256), 0, st, nedges, mask, ind, L,
                     (unsigned long long*)out);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

int launch_inner(xt_ctx* ctx, int64_t start, int64_t end, int64_t jstart, int64_t jend,
                 const int64_t* L, int64_t* out, hipStream_t st) {
  if (end <= start || jend <= jstart) return XT_OK;
  int64_t g = (end - start + 255) / 256;
  if (g > 65535) g = 65535;
  hipLaunchKernelGGL(k_inner, dim3((unsigned)g), dim3(256), 0, st, start, end, jstart, jend, L,
                     (unsigned long long*)out);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

int launch_add_total(xt_ctx* ctx, int64_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_add_total, dim3(1), dim3(64), 0, st,
                     (const uint64_t*)ctx->s.counters.as<uint64_t>(), (unsigned long long*)out);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

}  // namespace xt
