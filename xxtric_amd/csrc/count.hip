// count.hip -- intersection and reduction kernels over the oriented CSR (gfx950).
//
// Every oriented edge u->v closes one triangle per common out-neighbour w of u
// and v, so the count is the sum over oriented edges of |N+(u) & N+(v)|.
// Out-lists hold ascending ids.  d+(x) = |N+(x)|.
//
// Edges are binned by the longer of their two out-lists:
//  * light edges (both lists <= light_cut): one thread per edge walks the
//    shorter list and searches each id in the longer one (lower_bound restarted
//    after the previous hit) -- k_count_light, which also classifies the rest;
//  * heavy edges: the edge belongs to its endpoint x with the longer list (ties
//    to the source).  Each such x gets an item list -- the other endpoints y of
//    its heavy edges (k_fill_items).  k_count_groups, a persistent kernel,
//    draws (x, up to 256 items) tiles from a device work queue, stages N+(x)
//    once per tile in an LDS hash table, and streams the items' shorter lists
//    N+(y) through it with coalesced loads: each oriented edge reads only its
//    shorter list, the longer one is read once per tile.
// Reductions: per-thread uint64 -> wave64 shuffle tree -> LDS -> one 64-bit
// atomic per workgroup (per tile for the group kernel).
//
// Per-vertex credit (XT_PER_VERTEX): the triangle {a, b, w} belongs to
// min(a, b, w) (xxtric.hpp:165); hits credited to the edge's endpoints are
// summed in registers, hits credited to w use one atomic each.
#include "xt_device.hpp"
#include "xt_internal.hpp"

namespace xt {

constexpr int kGraphSlots = 64;        // per-tile LDS slots for per-graph batch totals
constexpr int kItemsPerTile = 256;     // items of one group tile (one per thread)
constexpr int kHashSlots = 8192;       // LDS hash capacity: staged lists up to 4096 ids
constexpr int kGroupThreads = 512;     // eight waves share one staged hash table

struct EdgeLists {
  int64_t s0, s1, l0, l1;  // shorter [s0,s1), longer [l0,l1) in otails
};

__device__ __forceinline__ EdgeLists edge_lists(int64_t a0, int64_t a1, int64_t b0, int64_t b1) {
  EdgeLists L;
  if (a1 - a0 <= b1 - b0) {
    L.s0 = a0; L.s1 = a1; L.l0 = b0; L.l1 = b1;
  } else {
    L.s0 = b0; L.s1 = b1; L.l0 = a0; L.l1 = a1;
  }
  return L;
}

// Rows of the oriented edges [eb, ee) staged in LDS when they fit.
struct TileRows {
  int64_t ulo, uhi, span;
  bool in_lds;
};

__device__ __forceinline__ TileRows tile_rows(const int64_t* __restrict__ ooff, int64_t nv,
                                              int64_t eb, int64_t ee, int64_t* s_off,
                                              int64_t* s_lohi) {
  if (threadIdx.x == 0) {
    const int64_t lo = row_of(ooff, 0, nv - 1, eb);
    s_lohi[0] = lo;
    s_lohi[1] = row_of(ooff, lo, nv - 1, ee - 1);
  }
  __syncthreads();
  TileRows r;
  r.ulo = s_lohi[0];
  r.uhi = s_lohi[1];
  r.span = r.uhi - r.ulo + 1;
  r.in_lds = r.span <= kOffCap;
  if (r.in_lds)
    for (int64_t k = threadIdx.x; k <= r.span; k += blockDim.x) s_off[k] = ooff[r.ulo + k];
  __syncthreads();
  return r;
}

__device__ __forceinline__ int64_t tile_row_of(const TileRows& r, const int64_t* s_off,
                                               const int64_t* __restrict__ ooff, int64_t e) {
  return r.in_lds ? r.ulo + lds_row_of(s_off, (int)r.span, e) : row_of(ooff, r.ulo, r.uhi, e);
}

__device__ __forceinline__ int64_t tile_off(const TileRows& r, const int64_t* s_off,
                                            const int64_t* __restrict__ ooff, int64_t u) {
  return r.in_lds ? s_off[u - r.ulo] : ooff[u];
}

// Ids of a thread's run of kPerThread oriented edges (two 16-byte loads when whole).
__device__ __forceinline__ void load_ids(const int32_t* __restrict__ ot, int64_t e0, int64_t ee,
                                         int32_t (&v)[kPerThread]) {
  if (e0 + kPerThread <= ee) {
    const int4* p = reinterpret_cast<const int4*>(ot + e0);
    const int4 a = p[0], b = p[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < kPerThread; j++) v[j] = e0 + j < ee ? ot[e0 + j] : 0;
  }
}

// ---------------------------------------------------------------------------
// Light edges counted inline; heavy edges counted per owning endpoint.
template <bool PV, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_count_light(
    int64_t nv, const int64_t* __restrict__ ooff, const int32_t* __restrict__ ot, int light_cut,
    uint64_t* __restrict__ counters, uint32_t* __restrict__ itcnt,
    unsigned long long* __restrict__ pv, const int32_t* __restrict__ gid,
    unsigned long long* __restrict__ gtotal) {
  __shared__ int64_t s_off[kOffCap + 1];
  __shared__ int64_t s_lohi[2];
  __shared__ uint64_t s_red[kBlock / kWave];
  __shared__ unsigned long long s_g[BATCH ? kGraphSlots : 1];
  const int64_t nep = ooff[nv];
  const int64_t eb = (int64_t)blockIdx.x * kChunk;
  if (eb >= nep) return;  // whole workgroup leaves together
  const int64_t ee = min(nep, eb + (int64_t)kChunk);
  if (BATCH && threadIdx.x < kGraphSlots) s_g[threadIdx.x] = 0;
  const TileRows r = tile_rows(ooff, nv, eb, ee, s_off, s_lohi);
  const int32_t g0 = BATCH ? gid[r.ulo] : 0;

  // Eight contiguous edges per thread: one row search, then a walk.
  uint64_t hits_all = 0;
  const int64_t e0 = eb + (int64_t)threadIdx.x * kPerThread;
  int32_t ids[kPerThread];
  if (e0 < ee) load_ids(ot, e0, ee, ids);
  int64_t u = e0 < ee ? tile_row_of(r, s_off, ooff, e0) : 0;
  int64_t a0 = e0 < ee ? tile_off(r, s_off, ooff, u) : 0;
  int64_t a1 = e0 < ee ? tile_off(r, s_off, ooff, u + 1) : 0;
  uint32_t own_run = 0;  // heavy edges of row u that u owns (source-owned)
#pragma unroll 1
  for (int j = 0; j < kPerThread; j++) {
    const int64_t e = e0 + j;
    if (e >= ee) break;
    while (e >= a1) {
      if (own_run) atomicAdd(&itcnt[u], own_run);
      own_run = 0;
      ++u;
      a0 = a1;
      a1 = tile_off(r, s_off, ooff, u + 1);
    }
    const int64_t v = ids[j];
    const int64_t b0 = ooff[v], b1 = ooff[v + 1];
    const int64_t du = a1 - a0, dv = b1 - b0;
    if (du > light_cut || dv > light_cut) {  // heavy: item of the longer list's owner
      if (du >= dv)
        own_run++;
      else
        atomicAdd(&itcnt[v], 1u);
      continue;
    }
    const EdgeLists L = edge_lists(a0, a1, b0, b1);
    if (L.s1 == L.s0) continue;
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
  if (own_run) atomicAdd(&itcnt[u], own_run);
  if constexpr (BATCH) {
    __syncthreads();
    if (threadIdx.x < kGraphSlots && s_g[threadIdx.x])
      atomicAdd(&gtotal[g0 + threadIdx.x], s_g[threadIdx.x]);
  } else {
    const uint64_t bsum = block_sum_u64<kBlock>(hits_all, s_red);
    if (threadIdx.x == 0 && bsum) atomicAdd((unsigned long long*)&counters[kTotal], (unsigned long long)bsum);
  }
}

// Writes the item lists: y goes to items[itoff[x] ..] of its owner x.  itcnt
// counts down to zero (slot order within a list is irrelevant).
__global__ __launch_bounds__(kBlock) void k_fill_items(int64_t nv, const int64_t* __restrict__ ooff,
                                                       const int32_t* __restrict__ ot, int light_cut,
                                                       const int64_t* __restrict__ itoff,
                                                       uint32_t* __restrict__ itcnt,
                                                       int32_t* __restrict__ items) {
  __shared__ int64_t s_off[kOffCap + 1];
  __shared__ int64_t s_lohi[2];
  const int64_t nep = ooff[nv];
  const int64_t eb = (int64_t)blockIdx.x * kChunk;
  if (eb >= nep) return;
  const int64_t ee = min(nep, eb + (int64_t)kChunk);
  const TileRows r = tile_rows(ooff, nv, eb, ee, s_off, s_lohi);
  const int64_t e0 = eb + (int64_t)threadIdx.x * kPerThread;
  if (e0 >= ee) return;
  int32_t ids[kPerThread];
  load_ids(ot, e0, ee, ids);
  int64_t u = tile_row_of(r, s_off, ooff, e0);
  int64_t a0 = tile_off(r, s_off, ooff, u), a1 = tile_off(r, s_off, ooff, u + 1);
  uint32_t own = 0;  // positions of this row's source-owned heavy edges
  // One atomic reserves a contiguous range for a run of source-owned items.
  auto flush = [&]() {
    if (!own) return;
    const uint32_t cnt = (uint32_t)__popc(own);
    const int64_t base = itoff[u] + (int64_t)(atomicSub(&itcnt[u], cnt) - cnt);
    uint32_t rank = 0;
#pragma unroll
    for (int jj = 0; jj < kPerThread; jj++)
      if ((own >> jj) & 1u) items[base + rank++] = ids[jj];
    own = 0;
  };
#pragma unroll
  for (int j = 0; j < kPerThread; j++) {
    const int64_t e = e0 + j;
    if (e < ee) {
      while (e >= a1) {
        flush();
        ++u;
        a0 = a1;
        a1 = tile_off(r, s_off, ooff, u + 1);
      }
      const int64_t v = ids[j];
      const int64_t du = a1 - a0, dv = ooff[v + 1] - ooff[v];
      if (du > light_cut || dv > light_cut) {
        if (du >= dv) {
          own |= 1u << j;
        } else {
          const uint32_t k = atomicSub(&itcnt[v], 1u) - 1u;
          items[itoff[v] + k] = (int32_t)u;
        }
      }
    }
  }
  flush();
}

__global__ __launch_bounds__(256) void k_tile_counts(int64_t nv, const int64_t* __restrict__ itoff,
                                                     uint32_t* __restrict__ tcnt) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += stride)
    tcnt[x] = (uint32_t)((itoff[x + 1] - itoff[x] + kItemsPerTile - 1) / kItemsPerTile);
}

__device__ __forceinline__ uint32_t hash_slot(int32_t w, int bits) {
  return ((uint32_t)w * 2654435761u) >> (32 - bits);
}

template <bool PV, bool BATCH>
__global__ __launch_bounds__(kGroupThreads) void k_count_groups(
    int64_t nv, const int64_t* __restrict__ ooff, const int32_t* __restrict__ ot,
    const int64_t* __restrict__ itoff, const int32_t* __restrict__ items,
    const int64_t* __restrict__ tileoff, uint64_t* __restrict__ counters,
    unsigned long long* __restrict__ pv, const int32_t* __restrict__ gid,
    unsigned long long* __restrict__ gtotal) {
  __shared__ int32_t s_hash[kHashSlots];
  __shared__ int64_t s_base[kItemsPerTile];
  __shared__ int32_t s_len[kItemsPerTile];
  __shared__ int32_t s_y[kItemsPerTile];
  __shared__ uint64_t s_red[kGroupThreads / kWave];
  __shared__ int64_t s_tile, s_x;
  __shared__ int s_next;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t ntiles = tileoff[nv];
  for (;;) {
    if (threadIdx.x == 0) {
      const int64_t t = (int64_t)atomicAdd((unsigned long long*)&counters[kTileQueue], 1ull);
      s_tile = t;
      s_x = t < ntiles ? row_of(tileoff, 0, nv - 1, t) : 0;
      s_next = 0;
    }
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile >= ntiles) break;  // uniform across the workgroup
    const int64_t x = s_x;
    const int64_t a0 = ooff[x], a1 = ooff[x + 1], dx = a1 - a0;
    const int64_t i0 = itoff[x] + (tile - tileoff[x]) * kItemsPerTile;
    const int64_t i1 = min(i0 + (int64_t)kItemsPerTile, itoff[x + 1]);
    const int nitems = (int)(i1 - i0);
    const bool hashed = dx <= kHashSlots / 2;
    // load factor <= 1/4 when the table fits (most probes miss: short probe runs)
    int bits = 6;
    while ((1ll << bits) < 4 * dx && (1 << bits) < kHashSlots) bits++;
    const uint32_t mask = (1u << bits) - 1u;
    if (hashed) {
      for (int k = threadIdx.x; k <= (int)mask; k += kGroupThreads) s_hash[k] = -1;
    }
    // items of this tile: bases and lengths of the streamed lists
    if ((int)threadIdx.x < nitems) {
      const int32_t y = items[i0 + threadIdx.x];
      const int64_t b = ooff[y];
      s_y[threadIdx.x] = y;
      s_base[threadIdx.x] = b;
      s_len[threadIdx.x] = (int32_t)(ooff[y + 1] - b);
    }
    if (hashed) {
      __syncthreads();  // table cleared before inserts
      for (int64_t k = threadIdx.x; k < dx; k += kGroupThreads) {
        const int32_t key = ot[a0 + k];
        uint32_t h = hash_slot(key, bits);
        while (atomicCAS(&s_hash[h], -1, key) != -1) h = (h + 1) & mask;
      }
    }
    __syncthreads();
    // Each wave takes whole items from an LDS counter and streams the item's
    // list with its 64 lanes: 16-byte loads (four ids per lane), four in flight
    // per lane, plus scalar loads for the unaligned head and the tail.
    uint64_t hits = 0, at_x = 0;
    int64_t yy = 0;
    auto probe = [&](int32_t w) {
      if (w < 0) return;
      bool found;
      if (hashed) {
        uint32_t h = hash_slot(w, bits);
        int32_t sl;
        while ((sl = s_hash[h]) != w && sl != -1) h = (h + 1) & mask;
        found = sl == w;
      } else {
        const int64_t p = lower_bound_dev<int32_t>(ot, a0, a1, w);
        found = p < a1 && ot[p] == w;
      }
      if (found) {
        hits++;
        if constexpr (PV) {
          int64_t m = x < yy ? x : yy;
          if ((int64_t)w < m) m = w;
          if (m == x)
            at_x++;
          else
            atomicAdd(&pv[m], 1ull);
        }
      }
    };
    for (;;) {
      int k = 0;
      if (lane == 0) k = atomicAdd(&s_next, 1);
      k = __shfl(k, 0, kWave);
      if (k >= nitems) break;  // uniform per wave
      const int64_t base = s_base[k];
      const int len = s_len[k];
      yy = s_y[k];
      const int32_t* __restrict__ lst = ot + base;
      int head = (int)((4 - (base & 3)) & 3);  // ids before the first 16-byte boundary
      if (head > len) head = len;
      if (lane < head) probe(lst[lane]);
      const int4* __restrict__ body = reinterpret_cast<const int4*>(lst + head);
      const int nbody = (len - head) >> 2;
      for (int i = lane; i < nbody; i += 4 * kWave) {
        int4 v[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
          v[q] = i + q * kWave < nbody ? body[i + q * kWave] : make_int4(-1, -1, -1, -1);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          probe(v[q].x);
          probe(v[q].y);
          probe(v[q].z);
          probe(v[q].w);
        }
      }
      const int tail0 = head + 4 * nbody;
      if (lane < len - tail0) probe(lst[tail0 + lane]);
    }
    __syncthreads();
    const uint64_t bh = block_sum_u64<kGroupThreads>(hits, s_red);
    if constexpr (PV) {
      __syncthreads();
      const uint64_t bx = block_sum_u64<kGroupThreads>(at_x, s_red);
      if (threadIdx.x == 0 && bx) atomicAdd(&pv[x], (unsigned long long)bx);
    }
    if (threadIdx.x == 0 && bh) {
      if constexpr (BATCH)
        atomicAdd(&gtotal[gid[x]], (unsigned long long)bh);
      else
        atomicAdd((unsigned long long*)&counters[kTotal], (unsigned long long)bh);
    }
    __syncthreads();  // LDS is reused by the next tile
  }
}

// Sums over oriented edges of min(d+(u), d+(v)) (all edges, and the heavy ones)
// and the number of heavy edges -- the algorithmic byte figures.
__global__ __launch_bounds__(kBlock) void k_alg_stats(int64_t nv, const int64_t* __restrict__ ooff,
                                                      const int32_t* __restrict__ ot, int light_cut,
                                                      uint64_t* __restrict__ counters) {
  __shared__ int64_t s_off[kOffCap + 1];
  __shared__ int64_t s_lohi[2];
  __shared__ uint64_t s_red[kBlock / kWave];
  const int64_t nep = ooff[nv];
  const int64_t eb = (int64_t)blockIdx.x * kChunk;
  if (eb >= nep) return;
  const int64_t ee = min(nep, eb + (int64_t)kChunk);
  const TileRows r = tile_rows(ooff, nv, eb, ee, s_off, s_lohi);
  uint64_t acc = 0, heavy = 0, nheavy = 0;
  for (int j = 0; j < kPerThread; j++) {
    const int64_t e = eb + (int64_t)j * kBlock + threadIdx.x;
    if (e >= ee) break;
    const int64_t u = tile_row_of(r, s_off, ooff, e);
    const int64_t v = ot[e];
    const int64_t du = tile_off(r, s_off, ooff, u + 1) - tile_off(r, s_off, ooff, u);
    const int64_t dv = ooff[v + 1] - ooff[v];
    const uint64_t m = (uint64_t)(du < dv ? du : dv);
    acc += m;
    if (du > light_cut || dv > light_cut) {
      heavy += m;
      nheavy++;
    }
  }
  uint64_t b = block_sum_u64<kBlock>(acc, s_red);
  if (threadIdx.x == 0 && b) atomicAdd((unsigned long long*)&counters[kStatMin], (unsigned long long)b);
  __syncthreads();
  b = block_sum_u64<kBlock>(heavy, s_red);
  if (threadIdx.x == 0 && b) atomicAdd((unsigned long long*)&counters[kStatMinHeavy], (unsigned long long)b);
  __syncthreads();
  b = block_sum_u64<kBlock>(nheavy, s_red);
  if (threadIdx.x == 0 && b) atomicAdd((unsigned long long*)&counters[kStatHeavyEdges], (unsigned long long)b);
}

int count_oriented(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, uint64_t* pv,
                   const int32_t* gid, uint64_t* gtotal, hipStream_t st, hipEvent_t* evs) {
  if (nv <= 0 || ne_bound <= 0) return XT_OK;
  const bool batch = gid != nullptr;
  const int64_t nchunks = (ne_bound + kChunk - 1) / kChunk;
  XT_TRY(ensure(ctx, s.itcnt, (size_t)nv * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.itoff, (size_t)(nv + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.items, (size_t)ne_bound * sizeof(int32_t)));
  XT_TRY(ensure(ctx, s.tileoff, (size_t)(nv + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.deg, (size_t)nv * sizeof(uint32_t)));  // reused as per-owner tile counts
  uint64_t* counters = s.counters.as<uint64_t>();
  XT_TRY(fill_u64(ctx, counters + kTileQueue, 1, 0, st));
  XT_TRY(fill_u32(ctx, s.itcnt.as<uint32_t>(), nv, 0u, st));
  const int64_t* ooff = s.ooff.as<int64_t>();
  const int32_t* ot = s.otails.as<int32_t>();
  unsigned long long* pvp = (unsigned long long*)pv;
  unsigned long long* gt = (unsigned long long*)gtotal;
  const int cut = ctx->small_cut;
#define XT_LIGHT(P, B)                                                                             \
  hipLaunchKernelGGL((k_count_light<P, B>), dim3((unsigned)nchunks), dim3(kBlock), 0, st, nv, ooff, \
                     ot, cut, counters, s.itcnt.as<uint32_t>(), pvp, gid, gt)
  if (pv && batch)
    XT_LIGHT(true, true);
  else if (pv)
    XT_LIGHT(true, false);
  else if (batch)
    XT_LIGHT(false, true);
  else
    XT_LIGHT(false, false);
#undef XT_LIGHT
  XT_HIP(ctx, hipGetLastError());
  if (evs) XT_HIP(ctx, hipEventRecord(evs[0], st));
  XT_TRY(scan_u32_to_i64(ctx, s, s.itcnt.as<uint32_t>(), nv, s.itoff.as<int64_t>(), st));
  hipLaunchKernelGGL(k_fill_items, dim3((unsigned)nchunks), dim3(kBlock), 0, st, nv, ooff, ot, cut,
                     (const int64_t*)s.itoff.as<int64_t>(), s.itcnt.as<uint32_t>(),
                     s.items.as<int32_t>());
  XT_HIP(ctx, hipGetLastError());
  int64_t g = (nv + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_tile_counts, dim3((unsigned)g), dim3(256), 0, st, nv,
                     (const int64_t*)s.itoff.as<int64_t>(), s.deg.as<uint32_t>());
  XT_HIP(ctx, hipGetLastError());
  XT_TRY(scan_u32_to_i64(ctx, s, s.deg.as<uint32_t>(), nv, s.tileoff.as<int64_t>(), st));
  if (evs) XT_HIP(ctx, hipEventRecord(evs[1], st));
  const unsigned groups = (unsigned)(ctx->num_cus * ctx->group_blocks);
#define XT_GROUPS(P, B)                                                                             \
  hipLaunchKernelGGL((k_count_groups<P, B>), dim3(groups), dim3(kGroupThreads), 0, st, nv, ooff, ot,       \
                     (const int64_t*)s.itoff.as<int64_t>(), (const int32_t*)s.items.as<int32_t>(), \
                     (const int64_t*)s.tileoff.as<int64_t>(), counters, pvp, gid, gt)
  if (pv && batch)
    XT_GROUPS(true, true);
  else if (pv)
    XT_GROUPS(true, false);
  else if (batch)
    XT_GROUPS(false, true);
  else
    XT_GROUPS(false, false);
#undef XT_GROUPS
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

int alg_stats(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, hipStream_t st) {
  if (nv <= 0 || ne_bound <= 0) return XT_OK;
  const int64_t nchunks = (ne_bound + kChunk - 1) / kChunk;
  uint64_t* counters = s.counters.as<uint64_t>();
  XT_TRY(fill_u64(ctx, counters + kStatMin, 1, 0, st));
  XT_TRY(fill_u64(ctx, counters + kStatMinHeavy, 2, 0, st));
  hipLaunchKernelGGL(k_alg_stats, dim3((unsigned)nchunks), dim3(kBlock), 0, st, nv,
                     (const int64_t*)s.ooff.as<int64_t>(), (const int32_t*)s.otails.as<int32_t>(),
                     ctx->small_cut, counters);
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
  XT_TRY(ensure(ctx, s.oflags, (size_t)(nchunks * kBlock)));
  XT_TRY(ensure(ctx, s.itcnt, (size_t)nvv * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.itoff, (size_t)(nvv + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.items, (size_t)nee * sizeof(int32_t)));
  XT_TRY(ensure(ctx, s.tileoff, (size_t)(nvv + 1) * sizeof(int64_t)));
  const size_t st = scan_tmp_bytes(nvv) > scan_tmp_bytes(nchunks) ? scan_tmp_bytes(nvv) : scan_tmp_bytes(nchunks);
  XT_TRY(ensure(ctx, s.scan_tmp, st));
  if (per_vertex) XT_TRY(ensure(ctx, s.pv, (size_t)nvv * sizeof(uint64_t)));
  return XT_OK;
}

void release(Scratch& s) {
  DevBuf* bufs[] = {&s.deg, &s.dplus, &s.tailcnt, &s.qscan, &s.ooff, &s.otails, &s.counters,
                    &s.oflags, &s.itcnt, &s.itoff, &s.items, &s.tileoff, &s.pv, &s.scan_tmp, &s.in_off,
                    &s.in_tails};
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
