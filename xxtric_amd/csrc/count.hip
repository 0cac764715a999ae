// count.hip -- intersection and reduction kernels over the oriented CSR (gfx950).
//
// Every oriented edge u->v closes one triangle per common out-neighbour w of u
// and v, so the count is the sum over oriented edges of |N+(u) & N+(v)|.
// Out-lists hold ascending ids.  d+(x) = |N+(x)|.
//
// Edges are binned by the longer of their two out-lists:
//  * light edges (both lists <= light_cut): intersected directly -- by the
//    whole wave, one edge at a time, each id of the shorter list searched in
//    the longer one, where light edges are sparse, or one lane per edge where
//    they are dense (blocks of eight ids in registers against 16-byte pieces
//    of the longer list);
//  * heavy edges: the edge belongs to its endpoint x with the longer list (ties
//    to the source).  Each such x gets an item list -- one 16-byte record
//    {offset, length, id} per other endpoint y of its heavy edges.
// Both happen in one pass over the symmetric input rows (k_light_items): entry
// (x, w) of row x is an out-edge or an in-edge of x, so every oriented edge is
// seen from both of its rows and is counted or listed from exactly one of them.
// The item lists are cut into tiles of up to 256 items (a tuning knob) with one
// descriptor each (k_tile_desc).  k_count_groups, a persistent kernel, draws
// runs of consecutive tiles (16 by default) from a device work queue, stages N+(x) in an LDS hash
// table of 4-slot buckets (kept for the next tile of the same owner), and streams the items' shorter lists N+(y) through it: the tile's
// lists are concatenated into one run of 16-byte chunks that all 512 threads
// read with coalesced loads, the next round of loads in flight while the
// current one is probed; compare results are lane masks, counted per wave on
// the scalar unit.  Each oriented edge reads only its shorter list; the longer
// one is read once per tile.
// Reductions: per-thread (and per-wave scalar) counters -> wave64 shuffle tree
// -> LDS -> one 64-bit atomic per workgroup (per tile for the group kernel).
//
// Per-vertex credit (XT_PER_VERTEX): the triangle {a, b, w} belongs to
// min(a, b, w) (xxtric.hpp:165); hits credited to the edge's endpoints are
// summed in registers, hits credited to w use one atomic each.
#include <type_traits>

#include "xt_device.hpp"
#include "xt_internal.hpp"

namespace xt {

constexpr int kGraphSlots = 64;        // per-tile LDS slots for per-graph batch totals
#ifndef XT_ITEMS_PER_TILE
#define XT_ITEMS_PER_TILE 256
#endif
constexpr int kItemsPerTile = XT_ITEMS_PER_TILE;  // most items of one group tile (LDS arrays)
constexpr int kMinItemsPerTile = 64;              // fewest a tiling may ask for (sizes the descriptors)
constexpr int kTabLog = 11;            // LDS hash: 2048 buckets of four ids (32 KB)
constexpr int kTabBuckets = 1 << kTabLog;
constexpr int kGroupThreads = 512;     // eight waves share one staged hash table
constexpr int kLutBlocks = 1024;       // chunk->item lookup for tiles of up to 64K chunks
constexpr int64_t kMaxCountEntries = (1ll << 33) - (1ll << 16);

// What a hit is credited to: the total only; per vertex, to the triangle's
// minimum vertex (xxtric.hpp:165); or, for an owned vertex range [lo, hi),
// to two totals -- triangles whose minimum vertex is owned, and of those the
// ones whose middle vertex is owned too (a rank of a 1-D vertex split counts
// the latter itself and forwards the rest as pairs, xxtric.hpp:189-204).
constexpr int kCreditTotal = 0;
constexpr int kCreditVertex = 1;
constexpr int kCreditRange = 2;

__device__ __forceinline__ void range_credit(int64_t a, int64_t b, int64_t c, int64_t lo, int64_t hi,
                                             uint32_t& owned, uint32_t& local) {
  const int64_t m = min(a, min(b, c)), M = max(a, max(b, c));
  const int64_t mid = a + b + c - m - M;
  if (m >= lo && m < hi) {
    owned++;
    if (mid < hi) local++;  // mid > m >= lo
  }
}

// One heavy edge seen from its owner x: the other endpoint y and its out-list.
struct alignas(16) Item {
  int64_t base;  // ooff[y]
  int32_t len;   // d+(y)
  int32_t y;
};

// Four ids at a 4-byte aligned address (one 16-byte load on gfx950).
typedef int v4i_a4 __attribute__((ext_vector_type(4), aligned(4)));
// The same in the global address space (a load through an address computed as
// an integer stays a global load, not a flat one).
typedef const __attribute__((address_space(1))) v4i_a4 gv4i_a4;
// Two consecutive offsets at an 8-byte aligned address (one 16-byte load).
typedef long long v2i64_a8 __attribute__((ext_vector_type(2), aligned(8)));

// One group tile: up to kItemsPerTile items of owner x and x's own out-list
// (32 bytes: the group kernel reads it as two int4 values in this field order).
struct alignas(16) TileDesc {
  int64_t i0;      // first item
  int64_t a0;      // ooff[x]
  int32_t x;
  int32_t nitems;
  int32_t dx;      // d+(x)
  int32_t pad;
};

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

// Light edges of one entry slot in a wave beyond which each lane intersects its
// own instead of the whole wave taking them one by one.
constexpr int kWaveWideLight = 4;

// |shorter ∩ longer| for two sorted lists of the oriented ids (both at most
// the light cut long), for plain counts.  Blocks of eight
// ids of the shorter list are held in registers and the longer list is walked
// in 16-byte pieces, each compared with the whole block (32 compares); a piece
// reaching past the block's largest id is kept for the next block.  About a
// dozen loads per edge, none dependent on a search.  Loads stay inside the ids
// plus their kOtailsPad padding: a piece starts before its list's end, and a
// block's second piece is read only where it does.
__device__ __forceinline__ uint32_t light_intersect(const int32_t* __restrict__ ot, const EdgeLists& L) {
  auto piece = [&](int64_t i) -> v4i_a4 {
    return *reinterpret_cast<const gv4i_a4*>(reinterpret_cast<uint64_t>(ot + i));
  };
  uint32_t h = 0;
  int64_t lc = L.l0;
  for (int64_t sc = L.s0; sc < L.s1; sc += 8) {
    const int64_t ns = L.s1 - sc;
    const v4i_a4 p = piece(sc);
    v4i_a4 q = {-1, -1, -1, -1};
    if (ns > 4) q = piece(sc + 4);
    const int32_t a[8] = {p.x,               ns > 1 ? p.y : -1, ns > 2 ? p.z : -1, ns > 3 ? p.w : -1,
                          ns > 4 ? q.x : -1, ns > 5 ? q.y : -1, ns > 6 ? q.z : -1, ns > 7 ? q.w : -1};
    const int32_t amax = max(max(max(a[0], a[1]), max(a[2], a[3])), max(max(a[4], a[5]), max(a[6], a[7])));
    while (lc < L.l1) {
      const int64_t nl = L.l1 - lc;
      const v4i_a4 t = piece(lc);
      const int32_t b[4] = {t.x, nl > 1 ? t.y : -2, nl > 2 ? t.z : -2, nl > 3 ? t.w : -2};
#pragma unroll
      for (int i = 0; i < 8; i++) h += (a[i] == b[0]) | (a[i] == b[1]) | (a[i] == b[2]) | (a[i] == b[3]);
      if (max(max(b[0], b[1]), max(b[2], b[3])) > amax) break;  // may meet the next block
      lc += 4;
    }
  }
  return h;
}

// The same for two lists of at most eight ids each (a light cut of 8 or
// less): all four 16-byte pieces are loaded at once and every pair of ids is
// compared, so the edge waits on one round of loads instead of a walk.
// Padding ids (-1 and -2) never match; loads stay inside each list's first
// eight ids or the ids that follow it (kOtailsPad at the end).
__device__ __forceinline__ uint32_t light_intersect8(const int32_t* __restrict__ ot, const EdgeLists& L) {
  auto piece = [&](int64_t i) -> v4i_a4 {
    return *reinterpret_cast<const gv4i_a4*>(reinterpret_cast<uint64_t>(ot + i));
  };
  const int64_t ns = L.s1 - L.s0, nl = L.l1 - L.l0;
  const v4i_a4 p = piece(L.s0), t = piece(L.l0);
  v4i_a4 q = {-1, -1, -1, -1}, u = {-2, -2, -2, -2};
  if (ns > 4) q = piece(L.s0 + 4);
  if (nl > 4) u = piece(L.l0 + 4);
  const int32_t a[8] = {p.x,               ns > 1 ? p.y : -1, ns > 2 ? p.z : -1, ns > 3 ? p.w : -1,
                        ns > 4 ? q.x : -1, ns > 5 ? q.y : -1, ns > 6 ? q.z : -1, ns > 7 ? q.w : -1};
  const int32_t b[8] = {t.x,               nl > 1 ? t.y : -2, nl > 2 ? t.z : -2, nl > 3 ? t.w : -2,
                        nl > 4 ? u.x : -2, nl > 5 ? u.y : -2, nl > 6 ? u.z : -2, nl > 7 ? u.w : -2};
  uint32_t h = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int k = 0; k < 8; k++) h += a[i] == b[k] ? 1u : 0u;
  return h;
}

// Lane l's value of v, in every lane (l uniform).
__device__ __forceinline__ int64_t readlane64(int64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), l);
  return (int64_t)(((uint64_t)hi << 32) | lo);
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

// Rows of input chunk c from the per-chunk row table, offsets staged in LDS.
__device__ __forceinline__ TileRows tile_rows_at(const int64_t* __restrict__ off,
                                                 const int64_t* __restrict__ crows, int64_t nv,
                                                 int64_t c, int64_t* s_off) {
  const RowSpan sp = chunk_span(crows, c, nv);
  TileRows r;
  r.ulo = sp.lo;
  r.uhi = sp.hi;
  r.span = r.uhi - r.ulo + 1;
  r.in_lds = r.span <= kOffCap;
  if (r.in_lds)
    for (int64_t k = threadIdx.x; k <= r.span; k += blockDim.x) s_off[k] = off[r.ulo + k];
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
// One pass over the symmetric input rows (2048 entries per workgroup, eight
// contiguous entries per thread).  Entry (x, w) of row x is the out-edge x->w
// when orientation kept it and the in-edge w->x otherwise (the input is
// symmetric).  With dx = d+(x), dw = d+(w):
//  * light edge (dx, dw <= light_cut), out-edge: intersected here -- the
//    thread walks the shorter list and searches each id in the longer one
//    (lower_bound restarted after the previous hit);
//  * heavy edge owned by x (out-edge with dx >= dw, or in-edge with dx > dw)
//    and dw > 0 (an empty list closes nothing): appended to x's item list,
//    which lives in items[off[x] ..] (row x has at most deg(x) items).  One
//    atomic per thread and row reserves the slots.
// Every oriented edge is therefore either counted here (once, from its source
// row) or listed once under its owner; in-edges and edges owned by the other
// endpoint are skipped.
//
// NARROW (|E+| < 2^32): one 8-byte gather per entry, vrec[w] = ooff[w] |
// d+(w) << 32, classifies the entry and places its item; exact out-degrees on
// both rows of an edge pick the same owner.  Otherwise a 2-byte gather of the
// saturated out-degree classifies every entry (both rows see the same
// saturated pair) and the owned and light entries gather their offsets.
// Six waves per SIMD for batches (dependent-load latency of small graphs'
// rows: 11.7 -> 11.4 ms on 4096 graphs), five otherwise (R-MAT s26's gathers:
// 42.6 ms at five, 46.3 at six).
template <typename T, bool NARROW, int CM, bool BATCH>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(BATCH && CM == kCreditTotal ? 6 : 5, 8))) void k_light_items(
    int64_t nv, int64_t ne, const int64_t* __restrict__ off, const T* __restrict__ tails,
    const uint8_t* __restrict__ oflags, const int64_t* __restrict__ ooff, const uint64_t* __restrict__ vrec,
    const uint4* __restrict__ vrange, const uint16_t* __restrict__ dp16, const int32_t* __restrict__ ot,
    const int64_t* __restrict__ crows, int light_cut, uint64_t* __restrict__ counters,
    uint32_t* __restrict__ itcnt, Item* __restrict__ items, unsigned long long* __restrict__ pv,
    const int32_t* __restrict__ gid, unsigned long long* __restrict__ gtotal, int64_t rlo, int64_t rhi) {
  constexpr bool PV = CM == kCreditVertex;
  // Owned-range counts of graphs below 2^32 entries are pruned: an edge whose
  // smaller endpoint a is below lo closes no owned triangle and is skipped; an
  // item streams only the ids w of its partner's list that can make the
  // triangle owned -- w >= lo when a < hi, lo <= w < hi when a >= hi -- a
  // contiguous piece of the sorted list whose bounds come with the record.
  constexpr bool PRUNE = NARROW && CM == kCreditRange;
  __shared__ int64_t s_off[kOffCap + 1];  // input offsets of the chunk's rows
  __shared__ uint64_t s_red[kBlock / kWave];
  __shared__ int64_t s_scan[kBlock / kWave];
  __shared__ uint16_t s_rank[kChunk + 1];    // owned entries before each entry of the chunk
  __shared__ uint32_t s_bb[2];               // reserved starts of the first and last row's items
  __shared__ unsigned long long s_g[BATCH ? kGraphSlots : 1];
  const int64_t cb = xcd_block((ne + kChunk - 1) / kChunk);
  const int64_t eb = cb * kChunk;
  if (eb >= ne) return;  // whole workgroup leaves together
  const int64_t ee = min(ne, eb + (int64_t)kChunk);
  if (BATCH && threadIdx.x < kGraphSlots) s_g[threadIdx.x] = 0;
  const TileRows r = tile_rows_at(off, crows, nv, cb, s_off);
  const int32_t g0 = BATCH ? gid[r.ulo] : 0;

  uint64_t hits_all = 0;
  uint32_t r_owned = 0, r_local = 0;  // owned-range credit
  const int64_t e0 = eb + (int64_t)threadIdx.x * kPerThread;
  int32_t ids[kPerThread];  // valid ids are below nv < 2^31
  uint32_t valid = 0, light = 0, own = 0;
  // Partner lists: base and length (32-bit each when NARROW).
  using LT = typename std::conditional<NARROW, uint32_t, int64_t>::type;
  LT gb[kPerThread], gl[kPerThread];
  uint32_t nlo[kPerThread], nhi[kPerThread];  // PRUNE: ids of the partner's list below lo / below hi
  int64_t u0 = 0;
  if (e0 < ee) {
    {
      int64_t raw[kPerThread];
      load_run<T>(tails, e0, ee, raw);
#pragma unroll
      for (int j = 0; j < kPerThread; j++) {
        const bool ok = e0 + j < ee && raw[j] >= 0 && raw[j] < nv;
        valid |= (uint32_t)ok << j;
        ids[j] = ok ? (int32_t)raw[j] : 0;
      }
    }
    const uint32_t fl = oflags[e0 / kPerThread];
    // 1. The eight neighbours' out-degrees (and, NARROW, their list offsets),
    //    eight gathers in flight.
    if constexpr (PRUNE) {
      uint4 rec[kPerThread];  // one 16-byte gather: offset, out-degree, ids below lo, ids below hi
#pragma unroll
      for (int j = 0; j < kPerThread; j++) rec[j] = vrange[ids[j]];
#pragma unroll
      for (int j = 0; j < kPerThread; j++) {
        gb[j] = rec[j].x;
        gl[j] = rec[j].y;
        nlo[j] = rec[j].z;
        nhi[j] = rec[j].w;
      }
    } else if constexpr (NARROW) {
      uint64_t rec[kPerThread];
#pragma unroll
      for (int j = 0; j < kPerThread; j++) rec[j] = vrec[ids[j]];
#pragma unroll
      for (int j = 0; j < kPerThread; j++) {
        gb[j] = (uint32_t)rec[j];
        gl[j] = (uint32_t)(rec[j] >> 32);
      }
    } else {
#pragma unroll
      for (int j = 0; j < kPerThread; j++) gl[j] = dp16[ids[j]];  // saturated until step 3
    }
    // 2. Classify along the rows: light out-edges are intersected here; heavy
    //    edges owned by this row are listed.
    u0 = tile_row_of(r, s_off, off, e0);
    {
      auto dplus_of = [&](int64_t uu) -> int64_t {
        if constexpr (NARROW && !PRUNE) return (int64_t)(vrec[uu] >> 32);  // one load: the row's own record
        const int64_t d = ooff[uu + 1] - ooff[uu];
        return NARROW ? d : (d < 65535 ? d : 65535);
      };
      int64_t u = u0;
      int64_t nxt = tile_off(r, s_off, off, u + 1);
      int64_t du_row = dplus_of(u);
#pragma unroll
      for (int j = 0; j < kPerThread; j++) {
        const int64_t e = e0 + j;
        if (e >= ee) break;
        while (e >= nxt) {
          ++u;
          nxt = tile_off(r, s_off, off, u + 1);
          du_row = dplus_of(u);
        }
        if (!((valid >> j) & 1u) || ids[j] == u) continue;  // out of range (unvalidated input) or self-loop
        const bool kept = (fl >> j) & 1u;
        const int64_t du = du_row, dv = (int64_t)gl[j];
        const int64_t a = u < (int64_t)ids[j] ? u : (int64_t)ids[j];
        if (du > light_cut || dv > light_cut) {  // heavy: listed under the longer list's owner
          if ((kept ? du >= dv : du > dv) && dv > 0) {  // an empty N+(v) closes nothing
            if constexpr (PRUNE) {
              if (a >= rlo) {
                const uint32_t s0 = nlo[j], s1 = a >= rhi ? nhi[j] : (uint32_t)dv;
                if (s1 > s0) {
                  gb[j] += s0;
                  gl[j] = s1 - s0;
                  own |= 1u << j;
                }
              }
            } else {
              own |= 1u << j;
            }
          }
        } else if (kept && dv > 0 && (!PRUNE || a >= rlo)) {  // light in-edges are counted from their source row
          light |= 1u << j;
        }
      }
    }
    // 3. Lists of the light and owned partners (about half the entries),
    //    gathered together (NARROW has them from step 1).
    if constexpr (!NARROW) {
#pragma unroll
      for (int j = 0; j < kPerThread; j++) {
        gb[j] = gl[j] = 0;
        if (((light | own) >> j) & 1u) {
          const v2i64_a8 t = *reinterpret_cast<const v2i64_a8*>(ooff + ids[j]);
          gb[j] = t.x;
          gl[j] = t.y - t.x;
        }
      }
    }
  }
  // 4. Item slots without per-row atomics: a block scan ranks the chunk's
  //    owned entries in entry order.  A row inside the chunk writes its items
  //    at off[x] + (rank within the row) and its count directly; only the
  //    chunk's first and last rows, which may continue in a neighbouring chunk,
  //    reserve their range with one atomic each.
  int64_t nown = 0;
  const int64_t pre = block_excl_scan_i64<kBlock>((int64_t)__popc(own), s_scan, &nown);
  if (e0 < ee) {
#pragma unroll
    for (int j = 0; j < kPerThread; j++)
      if (e0 + j < ee) s_rank[e0 + j - eb] = (uint16_t)(pre + __popc(own & ((1u << j) - 1u)));
  }
  if (threadIdx.x == 0) s_rank[ee - eb] = (uint16_t)nown;
  __syncthreads();
  if (threadIdx.x == 0) {
    auto reserve_row = [&](int64_t uu, int slot) {
      const int64_t ro = tile_off(r, s_off, off, uu), re = tile_off(r, s_off, off, uu + 1);
      s_bb[slot] = 0;
      if (ro < eb || re > ee) {  // continues into a neighbouring chunk
        const int cnt = s_rank[min(re, ee) - eb] - s_rank[max(ro, eb) - eb];
        if (cnt) s_bb[slot] = atomicAdd(&itcnt[uu], (uint32_t)cnt);
      }
    };
    reserve_row(r.ulo, 0);
    if (r.uhi != r.ulo) reserve_row(r.uhi, 1);
  }
  __syncthreads();
  {
    // 5. Along the rows again: write the owned items and intersect the light
    //    edges.  Where few lanes hold one (R-MAT: one light edge per few
    //    hundred entries), a light edge is intersected by the whole wave, one
    //    edge at a time: the lanes take the ids of the shorter list and search
    //    each in the longer one (both at most light_cut long), instead of one
    //    lane walking the list while the others wait.  Where many do (small
    //    sparse graphs), each lane intersects its own edge (light_intersect
    //    for plain counts).
    const bool live = e0 < ee;
    int64_t u = u0, cur = 0, nxt = 0, a0 = 0, a1 = 0;
    if (live) {
      cur = tile_off(r, s_off, off, u);
      nxt = tile_off(r, s_off, off, u + 1);
      if constexpr (NARROW && !PRUNE) {  // the row's own record: offset and out-degree in one load
        const uint64_t rec = vrec[u];
        a0 = (int64_t)(uint32_t)rec;
        a1 = a0 + (int64_t)(rec >> 32);
      } else {
        a0 = ooff[u];
        a1 = ooff[u + 1];
      }
    }
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int j = 0; j < kPerThread; j++) {
      const int64_t e = e0 + j;
      const bool act = live && e < ee;
      if (act) {
        while (e >= nxt) {
          ++u;
          cur = nxt;
          nxt = tile_off(r, s_off, off, u + 1);
          if constexpr (NARROW && !PRUNE) {
            const uint64_t rec = vrec[u];
            a0 = (int64_t)(uint32_t)rec;
            a1 = a0 + (int64_t)(rec >> 32);
          } else {
            a0 = a1;
            a1 = ooff[u + 1];
          }
        }
        const bool inner = cur >= eb && nxt <= ee;
        if (inner && e == cur) {  // the row's first entry publishes its item count
          const int cnt = s_rank[nxt - eb] - s_rank[cur - eb];
          if (cnt) itcnt[u] = (uint32_t)cnt;
        }
        if ((own >> j) & 1u) {
          const int64_t rank = s_rank[e - eb] - s_rank[max(cur, eb) - eb];
          const int64_t extra = inner ? 0 : (u == r.ulo ? (int64_t)s_bb[0] : (int64_t)s_bb[1]);
          items[cur + extra + rank] = Item{(int64_t)gb[j], (int32_t)gl[j], ids[j]};
        }
      }
      uint64_t pend = __builtin_amdgcn_ballot_w64(act && ((light >> j) & 1u));
      if (__popcll(pend) > kWaveWideLight) {  // uniform: many light edges, one per lane
        if (act && ((light >> j) & 1u)) {
          const int64_t v = ids[j];
          const EdgeLists L = edge_lists(a0, a1, (int64_t)gb[j], (int64_t)gb[j] + (int64_t)gl[j]);
          uint32_t h = 0;
          if constexpr (CM == kCreditTotal) {
            h = light_cut <= 8 ? light_intersect8(ot, L) : light_intersect(ot, L);  // uniform choice
          } else {
            // Credit modes need each common id: each id of the shorter list is
            // searched in the rest of the longer one (fewer registers than the
            // blocks, with the credit inside).
            const int64_t mmin = u < v ? u : v;
            int64_t pos = L.l0;
            for (int64_t k = L.s0; k < L.s1 && pos < L.l1; k++) {
              const int32_t w = ot[k];
              pos = lower_bound_dev<int32_t>(ot, pos, L.l1, w);  // restarted after the previous hit
              if (pos < L.l1 && ot[pos] == w) {
                h++;
                pos++;
                if constexpr (PV) {
                  atomicAdd(&pv[(int64_t)w < mmin ? (int64_t)w : mmin], 1ull);
                } else if constexpr (CM == kCreditRange) {
                  range_credit(u, v, w, rlo, rhi, r_owned, r_local);
                }
              }
            }
          }
          if constexpr (BATCH) {
            if (h) {
              const int32_t slot = gid[u] - g0;
              if (slot < kGraphSlots)
                atomicAdd(&s_g[slot], (unsigned long long)h);
              else
                atomicAdd(&gtotal[gid[u]], (unsigned long long)h);
            }
          } else {
            hits_all += h;
          }
        }
        pend = 0;
      }
      while (pend) {  // uniform
        const int src = __builtin_ctzll(pend);
        pend &= pend - 1;
        const int64_t su = readlane64(u, src), sv = __builtin_amdgcn_readlane(ids[j], src);
        const int64_t sb = readlane64((int64_t)gb[j], src);
        const EdgeLists L = edge_lists(readlane64(a0, src), readlane64(a1, src), sb,
                                       sb + readlane64((int64_t)gl[j], src));
        const int64_t mmin = su < sv ? su : sv;
        uint32_t h = 0;
        for (int64_t t = lane; t < L.s1 - L.s0; t += kWave) {
          const int32_t w = ot[L.s0 + t];
          const int64_t p = lower_bound_dev<int32_t>(ot, L.l0, L.l1, w);
          if (p < L.l1 && ot[p] == w) {
            h++;
            if constexpr (PV) {
              atomicAdd(&pv[(int64_t)w < mmin ? (int64_t)w : mmin], 1ull);
            } else if constexpr (CM == kCreditRange) {
              range_credit(su, sv, w, rlo, rhi, r_owned, r_local);
            }
          }
        }
        if constexpr (BATCH) {
          if (h) {
            const int32_t slot = gid[su] - g0;
            if (slot < kGraphSlots)
              atomicAdd(&s_g[slot], (unsigned long long)h);
            else
              atomicAdd(&gtotal[gid[su]], (unsigned long long)h);
          }
        } else {
          hits_all += h;
        }
      }
    }
  }
  if constexpr (BATCH) {
    __syncthreads();
    if (threadIdx.x < kGraphSlots && s_g[threadIdx.x])
      atomicAdd(&gtotal[g0 + threadIdx.x], s_g[threadIdx.x]);
  } else {
    const uint64_t bsum = block_sum_u64<kBlock>(hits_all, s_red);
    if (threadIdx.x == 0 && bsum) atomicAdd((unsigned long long*)&counters[kTotal], (unsigned long long)bsum);
  }
  if constexpr (CM == kCreditRange) {
    __syncthreads();
    const uint64_t bo = block_sum_u64<kBlock>((uint64_t)r_owned, s_red);
    if (threadIdx.x == 0 && bo) atomicAdd((unsigned long long*)&counters[kOwned], (unsigned long long)bo);
    __syncthreads();
    const uint64_t bl = block_sum_u64<kBlock>((uint64_t)r_local, s_red);
    if (threadIdx.x == 0 && bl) atomicAdd((unsigned long long*)&counters[kLocal], (unsigned long long)bl);
  }
  if (threadIdx.x == 0 && nown) atomicAdd((unsigned long long*)&counters[kHeavyItems], (unsigned long long)nown);
}

// Out-degrees saturated at 65535, from the oriented offsets.
__global__ __launch_bounds__(256) void k_dplus16(int64_t nv, const int64_t* __restrict__ ooff,
                                                 uint16_t* __restrict__ dp16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += stride) {
    const int64_t d = ooff[x + 1] - ooff[x];
    dp16[x] = (uint16_t)(d < 65535 ? d : 65535);
  }
}

// Per-vertex records ooff[x] | d+(x) << 32 (used when |E+| < 2^32, so both
// fit 32 bits): one 8-byte gather gives a neighbour's list offset and length.
__global__ __launch_bounds__(256) void k_vertex_records(int64_t nv, const int64_t* __restrict__ ooff,
                                                        uint64_t* __restrict__ vrec) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += stride) {
    const int64_t a = ooff[x], b = ooff[x + 1];
    vrec[x] = (uint64_t)(uint32_t)a | ((uint64_t)(uint32_t)(b - a) << 32);
  }
}

// Owned-range records {ooff[x], d+(x), ids of N+(x) below lo, ids below hi}
// (when |E+| < 2^32): the out-list is ascending, so the ids in [lo, hi) and
// those >= lo are contiguous pieces of it.
__global__ __launch_bounds__(256) void k_vertex_ranges(int64_t nv, const int64_t* __restrict__ ooff,
                                                       const int32_t* __restrict__ ot, int64_t lo, int64_t hi,
                                                       uint4* __restrict__ rec) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += stride) {
    const int64_t a = ooff[x], b = ooff[x + 1];
    const int64_t pl = lower_bound_dev<int32_t>(ot, a, b, (int32_t)lo);  // lo <= hi <= nv < 2^31
    const int64_t ph = lower_bound_dev<int32_t>(ot, pl, b, (int32_t)hi);
    rec[x] = make_uint4((uint32_t)a, (uint32_t)(b - a), (uint32_t)(pl - a), (uint32_t)(ph - a));
  }
}

// Wave tiles: an owner whose out-list has at most wave_owner ids (a cut of at
// most kWaveOwnerMax) is counted by k_count_waves, one wave per tile of at
// most kWaveItems items, instead of by the group kernel.
constexpr int kWaveItems = 64;

// Tiles of each owner: its items cut into tiles of ipt (<= kItemsPerTile) for
// the group kernel, counted at tcnt[x]; with wave tiles (wave_owner > 0) the
// owners with d+(x) <= wave_owner have tiles of kWaveItems counted at
// tcnt[nv + x] instead, so one scan over 2 nv counts puts every group tile
// before every wave tile.
__global__ __launch_bounds__(256) void k_tile_counts(int64_t nv, const uint32_t* __restrict__ itcnt,
                                                     const int64_t* __restrict__ ooff,
                                                     uint32_t* __restrict__ tcnt, int ipt, int wave_owner) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += stride) {
    const uint32_t it = itcnt[x];
    if (wave_owner > 0) {
      const bool wv = ooff[x + 1] - ooff[x] <= (int64_t)wave_owner;
      tcnt[x] = wv ? 0u : (it + (uint32_t)ipt - 1) / (uint32_t)ipt;
      tcnt[nv + x] = wv ? (it + (uint32_t)kWaveItems - 1) / (uint32_t)kWaveItems : 0u;
    } else {
      tcnt[x] = (it + (uint32_t)ipt - 1) / (uint32_t)ipt;
    }
  }
}

// One descriptor per tile; x's items start at off[x].  Group tiles are
// [tileoff[x], tileoff[x+1]), wave tiles [tileoff[nv+x], tileoff[nv+x+1]).
__global__ __launch_bounds__(256) void k_tile_desc(int64_t nv, const int64_t* __restrict__ off,
                                                   const int64_t* __restrict__ ooff,
                                                   const uint32_t* __restrict__ itcnt,
                                                   const int64_t* __restrict__ tileoff,
                                                   TileDesc* __restrict__ desc, int ipt_group, int wave_owner) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += stride) {
    const int64_t a0 = ooff[x];
    const int32_t dx = (int32_t)(ooff[x + 1] - a0);
    const bool wv = wave_owner > 0 && dx <= wave_owner;
    const int64_t t0 = wv ? tileoff[nv + x] : tileoff[x], t1 = wv ? tileoff[nv + x + 1] : tileoff[x + 1];
    if (t0 == t1) continue;
    const int ipt = wv ? kWaveItems : ipt_group;
    const int64_t ib = off[x], cnt = itcnt[x];
    for (int64_t t = t0; t < t1; t++) {
      const int64_t k = (t - t0) * ipt;
      const int64_t n = cnt - k < ipt ? cnt - k : ipt;
      desc[t] = TileDesc{ib + k, a0, (int32_t)x, (int32_t)n, dx, 0};
    }
  }
}

// LDS hash of N+(x): buckets of four ids; empty slots hold kEmpty = 2^31 - 1.
// A bucket that was full when a later key hashed to it has the top bit of its
// last slot set ("spilled"); that key went to the next bucket (linear probing
// by bucket).  A lookup reads its home bucket with one 16-byte LDS read and
// stops at the first bucket that holds the id or has not spilled.  Ids are
// below 2^31 - 1, so neither an empty nor a marked slot equals an id, and the
// padding id -1 matches nothing.
constexpr int32_t kEmpty = 0x7fffffff;

// Byte offset of w's home bucket, times 16.  MIX: the top 11 bits of
// w * 0x9E3779B1 (mod 2^32), a multiplicative hash -- ids that share their low
// bits (strided or bit-structured labels, such as an R-MAT graph's without a
// relabelling, whose ids are mostly zero bits) still spread over the buckets,
// and so do runs of consecutive ids.  Otherwise w mod 2048 (a shift and a
// mask): consecutive ids fall into distinct buckets, ids that agree in their
// low 11 bits share one and spill; on randomly labelled graphs it saves the
// multiply (R-MAT s26 with permuted labels: ~2 % of the count).
template <bool MIX>
__device__ __forceinline__ uint32_t bucket_addr(int32_t w) {
  if constexpr (MIX)
    return (((uint32_t)w * 0x9E3779B1u) >> (32 - kTabLog)) << 4;
  else
    return ((uint32_t)w << 4) & ((uint32_t)(kTabBuckets - 1) << 4);
}

__device__ __forceinline__ const int4& bucket_at(const int4* tab, uint32_t addr) {
  return *reinterpret_cast<const int4*>(reinterpret_cast<const char*>(tab) + addr);
}

// 0 when the bucket holds w (slot values xor w, minimum over the four slots).
__device__ __forceinline__ uint32_t bucket_miss(const int4 t, int32_t w) {
  const uint32_t x = (uint32_t)(t.x ^ w), y = (uint32_t)(t.y ^ w), z = (uint32_t)(t.z ^ w);
  const uint32_t q = (uint32_t)((t.w & 0x7fffffff) ^ w);
  return min(min(x, y), min(z, q));
}

__device__ __forceinline__ bool bucket_spilled(const int4 t) { return t.w < 0; }

// Inserts key; sets *spilled when a full bucket had to be passed.
template <bool MIX>
__device__ __forceinline__ void tab_insert(int4* tab, int32_t key, int* spilled) {
  uint32_t b = bucket_addr<MIX>(key) >> 4;
  for (;;) {
    int* s = reinterpret_cast<int*>(tab + b);
    if (atomicCAS(s, kEmpty, key) == kEmpty) return;
    if (atomicCAS(s + 1, kEmpty, key) == kEmpty) return;
    if (atomicCAS(s + 2, kEmpty, key) == kEmpty) return;
    if (atomicCAS(s + 3, kEmpty, key) == kEmpty) return;
    atomicOr(s + 3, (int)0x80000000);
    *spilled = 1;
    b = (b + 1) & (kTabBuckets - 1);
  }
}

// Continues a lookup past a spilled home bucket b.  Ends: at most half the
// slots are used, so some bucket is not full and therefore not spilled.
__device__ __forceinline__ bool tab_continue(const int4* tab, uint32_t b, int32_t w) {
  for (;;) {
    b = (b + 1) & (kTabBuckets - 1);
    const int4 t = tab[b];
    if (bucket_miss(t, w) == 0) return true;
    if (!bucket_spilled(t)) return false;
  }
}

// Slot index 4*bucket + s of w in bucket b (the four slots t), or -1.
__device__ __forceinline__ int bucket_slot(const int4 t, uint32_t b, int32_t w) {
  const int s = t.x == w ? 0 : t.y == w ? 1 : t.z == w ? 2 : (t.w & 0x7fffffff) == w ? 3 : -1;
  return s < 0 ? -1 : (int)(4 * b) + s;
}

// tab_continue returning the slot index of w, or -1.
__device__ __forceinline__ int tab_continue_slot(const int4* tab, uint32_t b, int32_t w) {
  for (;;) {
    b = (b + 1) & (kTabBuckets - 1);
    const int4 t = tab[b];
    const int s = bucket_slot(t, b, w);
    if (s >= 0) return s;
    if (!bucket_spilled(t)) return -1;
  }
}

template <int CM, bool BATCH, int RC, int WPE, bool MIX>
__global__ __launch_bounds__(kGroupThreads) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_count_groups(
    const int32_t* __restrict__ ot, const Item* __restrict__ items, const TileDesc* __restrict__ tdesc,
    const int64_t* __restrict__ ntiles_p, uint64_t* __restrict__ counters,
    unsigned long long* __restrict__ pv, const int32_t* __restrict__ gid,
    unsigned long long* __restrict__ gtotal, int64_t rlo, int64_t rhi, uint32_t tile_run) {
  constexpr bool PV = CM == kCreditVertex;
  __shared__ int4 s_tab[kTabBuckets];
  // Per item k: {adj (64-bit), end, y}.  Chunk c of item k (16 bytes) starts at
  // byte address adj + 16*c and holds end - 4*c ids of the list (mod 2^32;
  // <= 0 past the item's last chunk).
  __shared__ int4 s_rec[kItemsPerTile];
  __shared__ int32_t s_pre[kItemsPerTile + 1];   // exclusive scan of the items' chunk counts
  __shared__ uint16_t s_lut[kLutBlocks];          // item holding chunk 64*b
  __shared__ unsigned long long s_starts[kLutBlocks];  // items starting inside block b (bit = offset)
  __shared__ int64_t s_scan[kGroupThreads / kWave];
  __shared__ uint64_t s_red[kGroupThreads / kWave];
  __shared__ int4 s_desc[2][2];  // TileDesc of the current and of the next tile (ping-pong)
  __shared__ uint32_t s_tile;
  __shared__ int s_spilled;  // some bucket of the current table spilled
  // Per-vertex credit, gathered in LDS before it goes to HBM: hits whose
  // smallest vertex is the table id in slot i add to the 16-bit half i & 1 of
  // s_wcnt[i >> 1] (kept for up to kFlushTiles tiles of one table, so below
  // kFlushTiles * kItemsPerTile, as an item hits each id at most once); hits
  // whose smallest vertex is item k's partner add to s_ycnt[k] (one tile).
  constexpr int kFlushTiles = 16;
  static_assert(kFlushTiles * kItemsPerTile < 65536, "16-bit slot counters");
  __shared__ uint32_t s_wcnt[PV ? 2 * kTabBuckets : 1];
  __shared__ uint32_t s_ycnt[PV ? kItemsPerTile : 1];
  constexpr int kT = kGroupThreads;
  constexpr int kJ = (kItemsPerTile + kT - 1) / kT;  // items per thread in the tile set-up
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);  // uniform per wave
  // Tile indices fit 32 bits: there are fewer than nv + ne/64 + 1 tiles, with
  // nv < 2^31 and ne < 2^33 (checked on the host), so the queue, which overshoots
  // ntiles by at most 3 runs of at most 64 tiles per workgroup, cannot wrap.
  const uint32_t ntiles = (uint32_t)*ntiles_p;
  unsigned int* queue = reinterpret_cast<unsigned int*>(&counters[kTileQueue]);

  // Work queue, fetched ahead by wave 0: while tile T runs, the descriptor of
  // T+1 is loaded and the index of T+2 is drawn.
  // Tiles are drawn in runs of tile_run consecutive indices: an owner's tiles
  // are consecutive, so a workgroup often keeps the owner's table for the
  // next tile.
  uint32_t run_next = 0, run_end = 0;  // wave 0: tiles drawn, not yet taken
  auto draw = [&]() -> uint32_t {
    if (run_next >= run_end) {
      uint32_t v = 0;
      if (lane == 0) v = atomicAdd(queue, tile_run);
      run_next = __builtin_amdgcn_readfirstlane(v);
      run_end = run_next + tile_run;
    }
    return run_next++;
  };
  int64_t table_a0 = -1;  // owner list staged in s_tab (its oriented offset), -1 none
  // Descriptors are copied global -> LDS by wave 0's lanes 0-1 with one 16-byte
  // LDS-DMA load each (no registers held across the tile).
  const int4* __restrict__ tdesc4 = reinterpret_cast<const int4*>(tdesc);
  auto fetch_desc = [&](uint32_t t, int slot) {
    if (lane < 2) __builtin_amdgcn_global_load_lds(tdesc4 + 2 * (size_t)t + lane, &s_desc[slot][0], 16, 0, 0);
  };
  uint32_t t_nx = 0, t_nx2 = 0;
  int par = 0;
  int table_tiles = 0;  // tiles counted into s_wcnt since its last flush
  if constexpr (PV) {
    for (int i = tid; i < 2 * kTabBuckets; i += kT) s_wcnt[i] = 0;
    for (int i = tid; i < kItemsPerTile; i += kT) s_ycnt[i] = 0;
  }
  // Adds the slot counters of the table in s_tab to pv and clears them.  Each
  // thread takes whole buckets, the same ones it clears and fills in the set-up.
  auto flush_slots = [&]() {
    if constexpr (PV) {
      for (int b = tid; b < kTabBuckets; b += kT) {
        const uint32_t c01 = s_wcnt[2 * b], c23 = s_wcnt[2 * b + 1];
        if ((c01 | c23) == 0) continue;
        const int4 t = s_tab[b];
        if (c01 & 0xffffu) atomicAdd(&pv[t.x], (unsigned long long)(c01 & 0xffffu));
        if (c01 >> 16) atomicAdd(&pv[t.y], (unsigned long long)(c01 >> 16));
        if (c23 & 0xffffu) atomicAdd(&pv[t.z], (unsigned long long)(c23 & 0xffffu));
        if (c23 >> 16) atomicAdd(&pv[t.w & 0x7fffffff], (unsigned long long)(c23 >> 16));
        s_wcnt[2 * b] = 0;
        s_wcnt[2 * b + 1] = 0;
      }
    }
  };
  if (wave == 0) {
    const uint32_t t = draw();
    if (lane == 0) s_tile = t;
    if (t < ntiles) fetch_desc(t, 0);
    t_nx = draw();
    __builtin_amdgcn_s_waitcnt(0);  // the LDS-DMA copy has landed
  }
  __syncthreads();
  for (;;) {
    const uint32_t tile = s_tile;
    if (tile >= ntiles) break;  // uniform across the workgroup
    TileDesc d;
    {
      const int4 u0 = s_desc[par][0], u1 = s_desc[par][1];
      d.i0 = (int64_t)(((uint64_t)(uint32_t)u0.y << 32) | (uint32_t)u0.x);
      d.a0 = (int64_t)(((uint64_t)(uint32_t)u0.w << 32) | (uint32_t)u0.z);
      d.x = u1.x;
      d.nitems = u1.y;
      d.dx = u1.z;
    }
    if (wave == 0) {
      t_nx2 = t_nx;
      if (t_nx < ntiles) {
        fetch_desc(t_nx, par ^ 1);  // the other slot was last read before the previous barrier
        t_nx2 = draw();
      }
    }
    const int64_t x = d.x;
    const int nitems = d.nitems;
    const int64_t i0 = d.i0, a0 = d.a0, dx = d.dx;
    const bool hashed = dx <= 2 * kTabBuckets;  // at most two ids per bucket on average
    const bool reuse = hashed && a0 == table_a0;  // the previous tile's owner: table in place

    // Items k = tid*kJ + j: each list is streamed as ceil(len/4) 16-byte chunks
    // from its first id (dword-aligned loads; the last chunk is masked).
    int32_t nc[kJ];
    Item it[kJ];
    int32_t nchunks = 0;
#pragma unroll
    for (int j = 0; j < kJ; j++) {
      const int k = tid * kJ + j;
      nc[j] = 0;
      if (k < nitems) {
        it[j] = items[i0 + k];
        nc[j] = (it[j].len + 3) >> 2;
        nchunks += nc[j];
      }
    }
    if constexpr (PV) {
      // The previous table's counters go out before it is replaced, and at
      // least every kFlushTiles tiles (16-bit counters).
      if (hashed && (!reuse || table_tiles == kFlushTiles)) {
        flush_slots();
        table_tiles = 0;
      }
      if (hashed) table_tiles++;
    }
    if (hashed && !reuse) {
      for (int b = tid; b < kTabBuckets; b += kT) s_tab[b] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
      if (tid == 0) s_spilled = 0;
    }
    int64_t nch_all = 0;
    int64_t pre = block_excl_scan_i64<kT>(nchunks, s_scan, &nch_all);
#pragma unroll
    for (int j = 0; j < kJ; j++) {
      const int k = tid * kJ + j;
      if (k < nitems) {
        s_pre[k] = (int32_t)pre;
        const uint64_t adj = (uint64_t)(uintptr_t)(ot + it[j].base) - 16ull * (uint64_t)pre;
        s_rec[k] = make_int4((int32_t)(uint32_t)adj, (int32_t)(uint32_t)(adj >> 32),
                             (int32_t)((uint32_t)it[j].len + 4u * (uint32_t)pre), it[j].y);
      }
      pre += nc[j];
    }
    if (tid == 0) s_pre[nitems] = (int32_t)nch_all;
    // Entries past the last item never compare <= a chunk index, so the
    // binary search below needs no bounds test.
    for (int k = nitems + 1 + tid; k <= kItemsPerTile; k += kT) s_pre[k] = INT32_MAX;
    const int nchunk = (int)nch_all;
    const int nblk = (nchunk + kWave - 1) / kWave;
    const bool lut = nblk <= kLutBlocks;  // uniform
    if (lut)
      for (int b = tid; b < nblk; b += kT) s_starts[b] = 0ull;
    __syncthreads();  // table cleared, item table and chunk scan published

    // Chunk -> item.  Chunks are taken 64 at a time, one per lane, from
    // 64-aligned blocks, so a wave's lanes share one block b: the item is
    // s_lut[b] (the item holding the block's first chunk) plus the items that
    // start inside the block at or before the lane.  Larger tiles search s_pre.
    auto search = [&](int c) {
      int kk = 0;
#pragma unroll
      for (int step = kItemsPerTile / 2; step > 0; step >>= 1)
        kk = s_pre[kk + step] <= c ? kk + step : kk;
      return kk;
    };
    if (lut) {
      for (int b = tid; b < nblk; b += kT) s_lut[b] = (uint16_t)search(b * kWave);
#pragma unroll
      for (int j = 0; j < kJ; j++) {
        const int k = tid * kJ + j;
        if (k < nitems) {
          const uint32_t c = (uint32_t)s_pre[k];
          if (c & (kWave - 1)) atomicOr(&s_starts[c / kWave], 1ull << (c & (kWave - 1)));
        }
      }
    }
    // Bits 1..lane: item starts inside the block up to and including this lane.
    const unsigned long long upto = (2ull << lane) - 2ull;
    // Chunks go out in 64-chunk blocks, one chunk per lane: in round r, wave w
    // takes blocks (r*RC + q)*8 + w for q < RC, so even a small tile spreads
    // over all eight waves.  A wave whose block lies past the end still issues
    // its load (the last chunk again, counted as no ids) so that every round
    // has the same loads in flight, but skips the lookup and the probes.
    constexpr int kWaves = kT / kWave;
    auto blk = [&](int round, int q) { return (round * RC + q) * kWaves + wave; };
    auto locate = [&](int round, int (&kq)[RC], int4 (&vq)[RC], int (&rq)[RC]) {
#pragma unroll
      for (int q = 0; q < RC; q++) {
        const int b = blk(round, q);
        const int c = b * kWave + lane;
        const int cc = c < nchunk ? c : nchunk - 1;  // past the end: the last chunk again
        int kk = nitems - 1;
        if (b < nblk) kk = lut ? (int)s_lut[b] + __popcll(s_starts[b] & upto) : search(cc);  // uniform
        const int4 rec = s_rec[kk];
        kq[q] = kk;
        rq[q] = (int32_t)((uint32_t)rec.z - 4u * (uint32_t)c);  // <= 0 past the end
        const uint64_t adj = ((uint64_t)(uint32_t)rec.y << 32) | (uint32_t)rec.x;
        const v4i_a4 t = *reinterpret_cast<const gv4i_a4*>(adj + 16ull * (uint32_t)cc);
        vq[q] = make_int4(t.x, t.y, t.z, t.w);
      }
    };
    __syncthreads();  // chunk lookup built
    constexpr int kRound = kT * RC;  // chunks per round
    int kA[RC], rA[RC], kB[RC], rB[RC];
    int4 vA[RC], vB[RC];
    if (nchunk > 0) locate(0, kA, vA, rA);  // first loads fly while the table is filled
    if (hashed && !reuse)
      for (int64_t k = tid; k < dx; k += kT) tab_insert<MIX>(s_tab, ot[a0 + k], &s_spilled);
    __syncthreads();  // table complete
    table_a0 = hashed ? a0 : -1;

    const bool spl = s_spilled != 0;  // uniform: read after the table barrier
    uint32_t hits = 0, at_x = 0;  // below 2^32: a tile streams fewer than 2^33 ids over 512 threads
    uint32_t r_owned = 0, r_local = 0;  // owned-range credit
    // Hits counted per wave (the same value in every lane): below 2^32, since the
    // waves take the tile's chunks in turns and it streams fewer than 2^33 ids.
    uint32_t hits_w = 0;
    uint32_t r_owned_w = 0, r_local_w = 0;  // owned-range credit counted per wave
    // A hit, with its per-vertex or owned-range credit.  slot: w's slot in the
    // LDS table, -1 when the owner list is searched in HBM instead.
    auto credit = [&](int32_t w, int k, int slot) {
      hits++;
      if constexpr (PV) {
        const int64_t yy = s_rec[k].w;  // x, y and w are three distinct vertices
        if (x < yy && x < (int64_t)w)
          at_x++;
        else if (yy < (int64_t)w)
          atomicAdd(&s_ycnt[k], 1u);
        else if (slot >= 0)
          atomicAdd(&s_wcnt[slot >> 1], 1u << (16 * (slot & 1)));
        else
          atomicAdd(&pv[w], 1ull);
      } else if constexpr (CM == kCreditRange) {
        range_credit(x, (int64_t)s_rec[k].w, (int64_t)w, rlo, rhi, r_owned, r_local);
      }
    };
    // Per-vertex probe: membership of the ids of one chunk of item k (r of its
    // four ids lie inside the list) in the LDS hash: four bucket reads and raw
    // compares, as in the plain probe; the slot index is formed from the
    // compares only for a hit credited to its table id.  Each triangle
    // {x, y, w} goes to its smallest vertex: x (a register count), y (the
    // chunk's partner count, one LDS add per chunk) or w (its table slot's
    // counter, one LDS add per such hit).
    auto probe_pv = [&](const int4 v, int k, int r) {
      const int32_t w[4] = {v.x, v.y, v.z, v.w};
      int4 t[4];
#pragma unroll
      for (int j = 0; j < 4; j++) t[j] = bucket_at(s_tab, bucket_addr<MIX>(w[j]));
      const int32_t x32 = (int32_t)x;  // ids are below 2^31
      const int32_t y = s_rec[k].w;
      const int32_t lo = min(x32, y);
      uint32_t nlo = 0;  // hits whose smallest vertex is min(x, y)
      // One hit of w at slot index `slot`: counted, and credited to w when w is
      // the smallest of the three.
      auto hit = [&](int32_t wj, int slot) {
        hits++;
        if (wj < lo)
          atomicAdd(&s_wcnt[slot >> 1], 1u << (16 * (slot & 1)));
        else
          nlo++;
      };
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const bool e0 = t[j].x == w[j], e1 = t[j].y == w[j], e2 = t[j].z == w[j], e3 = t[j].w == w[j];
        if (j < r && (e0 | e1 | e2 | e3))
          hit(w[j], (int)(bucket_addr<MIX>(w[j]) >> 2) + ((e1 | e3) ? 1 : 0) + ((e2 | e3) ? 2 : 0));
      }
      // Slot 3 of a spilled bucket carries the mark, so the raw compare missed an
      // id there; lanes with a spilled bucket re-test it and walk on (only
      // tables with a spilled bucket can have one; uniform test).
      if (spl && __builtin_amdgcn_ballot_w64((t[0].w | t[1].w | t[2].w | t[3].w) < 0)) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (j >= r || t[j].w >= 0) continue;
          if (t[j].x == w[j] || t[j].y == w[j] || t[j].z == w[j]) continue;  // counted above
          int slot = -1;
          if ((t[j].w & 0x7fffffff) == w[j])
            slot = (int)(bucket_addr<MIX>(w[j]) >> 2) + 3;
          else
            slot = tab_continue_slot(s_tab, bucket_addr<MIX>(w[j]) >> 4, w[j]);
          if (slot >= 0) hit(w[j], slot);
        }
      }
      if (x32 < y)
        at_x += nlo;
      else if (nlo)
        atomicAdd(&s_ycnt[k], nlo);
    };
    // Counting without per-vertex credit: each id is compared with the four
    // slots of its home bucket as raw words; the compare results are lane
    // masks whose population counts are summed per wave.  Slot 3 of a spilled
    // bucket carries the mark, so an id there compares unequal -- like every
    // other miss in a spilled bucket it goes to the continuation, which tests
    // the bucket again with the mark removed and walks on.
    // With owned-range credit the same masks are split further: the triangle
    // {x, y, w} of a hit has minimum min(a, w) and middle max(a, min(b, w)),
    // a = min(x, y), b = max(x, y).
    auto probe_count = [&](const int4 v, int k, int r) {
      const int32_t w[4] = {v.x, v.y, v.z, v.w};
      int4 t[4];
#pragma unroll
      for (int j = 0; j < 4; j++) t[j] = bucket_at(s_tab, bucket_addr<MIX>(w[j]));
      int32_t ya = 0, yb = 0;
      if constexpr (CM == kCreditRange) {
        const int32_t y = s_rec[k].w;
        ya = min((int32_t)x, y);
        yb = max((int32_t)x, y);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint64_t hit = (__builtin_amdgcn_ballot_w64(t[j].x == w[j]) | __builtin_amdgcn_ballot_w64(t[j].y == w[j]) |
                              __builtin_amdgcn_ballot_w64(t[j].z == w[j]) | __builtin_amdgcn_ballot_w64(t[j].w == w[j])) &
                             __builtin_amdgcn_ballot_w64(j < r);
        hits_w += (uint32_t)__popcll(hit);
        if constexpr (CM == kCreditRange) {
          const int32_t m = min(ya, w[j]), mid = max(ya, min(yb, w[j]));
          const uint64_t own =
              hit & __builtin_amdgcn_ballot_w64((uint32_t)(m - (int32_t)rlo) < (uint32_t)(rhi - rlo));
          r_owned_w += (uint32_t)__popcll(own);
          r_local_w += (uint32_t)__popcll(own & __builtin_amdgcn_ballot_w64(mid < (int32_t)rhi));
        }
      }
      // A lane whose four buckets include a spilled one takes the continuation
      // (only tables with a spilled bucket can have one; uniform test).
      if (spl && __builtin_amdgcn_ballot_w64((t[0].w | t[1].w | t[2].w | t[3].w) < 0)) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (j >= r || t[j].w >= 0) continue;
          bool h = (t[j].w & 0x7fffffff) == w[j];  // in slot 3 of a spilled bucket: the raw compare missed it
          if (!h) h = bucket_miss(t[j], w[j]) != 0 && tab_continue(s_tab, bucket_addr<MIX>(w[j]) >> 4, w[j]);
          if (h) {
            hits++;
            if constexpr (CM == kCreditRange) range_credit(ya, yb, w[j], rlo, rhi, r_owned, r_local);
          }
        }
      }
    };
    // Lists too long for the LDS hash: binary search in N+(x) in global memory.
    auto probe_search = [&](const int4 v, int k, int r) {
      const int32_t w[4] = {r > 0 ? v.x : -1, r > 1 ? v.y : -1, r > 2 ? v.z : -1, r > 3 ? v.w : -1};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (w[j] < 0) continue;
        const int64_t p = lower_bound_dev<int32_t>(ot, a0, a0 + dx, w[j]);
        if (p < a0 + dx && ot[p] == w[j]) credit(w[j], k, -1);
      }
    };

    auto probe = [&](const int4 v, int k, int r) {
      if constexpr (PV)
        probe_pv(v, k, r);
      else
        probe_count(v, k, r);
    };
    // Two register buffers alternate by hand (no loop-carried copies) and the
    // next round is always issued (past the end it re-reads the last chunk,
    // masked), so every round waits only for its own loads while the next
    // round's are in flight.
    if (hashed && nchunk > 0) {
      for (int r0 = 0;; r0 += 2) {
        locate(r0 + 1, kB, vB, rB);
#pragma unroll
        for (int q = 0; q < RC; q++)
          if (blk(r0, q) < nblk) probe(vA[q], kA[q], rA[q]);  // uniform per wave
        if ((r0 + 1) * kRound >= nchunk) break;  // uniform
        locate(r0 + 2, kA, vA, rA);
#pragma unroll
        for (int q = 0; q < RC; q++)
          if (blk(r0 + 1, q) < nblk) probe(vB[q], kB[q], rB[q]);
        if ((r0 + 2) * kRound >= nchunk) break;  // uniform
      }
    } else if (!hashed) {
      for (int r0 = 0; r0 * kRound < nchunk; r0++) {
        if (r0 > 0) locate(r0, kA, vA, rA);
#pragma unroll
        for (int q = 0; q < RC; q++)
          if (blk(r0, q) < nblk) probe_search(vA[q], kA[q], rA[q]);
      }
    }

    const uint64_t bh = block_sum_u64<kT>((uint64_t)hits + (lane == 0 ? (uint64_t)hits_w : 0ull), s_red);
    if constexpr (PV) {
      __syncthreads();
      const uint64_t bx = block_sum_u64<kT>((uint64_t)at_x, s_red);
      if (tid == 0 && bx) atomicAdd(&pv[x], (unsigned long long)bx);
      // Partner credit of the tile's items (the block sums above ended every probe).
      for (int k = tid; k < nitems; k += kT) {
        const uint32_t c = s_ycnt[k];
        if (c) {
          atomicAdd(&pv[s_rec[k].w], (unsigned long long)c);
          s_ycnt[k] = 0;
        }
      }
    }
    if constexpr (CM == kCreditRange) {
      __syncthreads();
      const uint64_t bo = block_sum_u64<kT>((uint64_t)r_owned + (lane == 0 ? (uint64_t)r_owned_w : 0ull), s_red);
      if (tid == 0 && bo) atomicAdd((unsigned long long*)&counters[kOwned], (unsigned long long)bo);
      __syncthreads();
      const uint64_t bl = block_sum_u64<kT>((uint64_t)r_local + (lane == 0 ? (uint64_t)r_local_w : 0ull), s_red);
      if (tid == 0 && bl) atomicAdd((unsigned long long*)&counters[kLocal], (unsigned long long)bl);
    }
    if (tid == 0 && bh) {
      if constexpr (BATCH)
        atomicAdd(&gtotal[gid[x]], (unsigned long long)bh);
      else
        atomicAdd((unsigned long long*)&counters[kTotal], (unsigned long long)bh);
    }
    if (wave == 0) {
      __builtin_amdgcn_s_waitcnt(0);  // the next descriptor's LDS-DMA copy has landed
      if (lane == 0) s_tile = t_nx;
    }
    t_nx = t_nx2;
    par ^= 1;
    __syncthreads();  // next tile published; LDS free for reuse
  }
  flush_slots();  // the last table's counters (the loop ends on a barrier)
}

// ---------------------------------------------------------------------------
// Wave tiles (plain counts).  An owner x with a short out-list gets one wave,
// not a 512-thread workgroup with a 2048-bucket table: the wave hashes N+(x)
// (at most kWaveOwnerMax ids) into its own LDS table of 64, 128 or 256
// buckets (1, 2 or 4 KB: at most half its slots filled) and
// streams its tile's items -- at most 64, one per lane -- through it.  The
// items' lists are concatenated into 16-byte chunks and dealt out 64 at a
// time, one per lane, so lanes never wait on a longer list than their own
// chunk.  Small sparse graphs (batches) have almost only such owners, and a
// workgroup-wide tile per owner cost more in set-up than its lists.
// Chunk c of a round belongs to the item k with pre[k] <= c < pre[k+1]
// (pre: exclusive scan of the items' chunk counts, lane k holding pre[k]),
// found by a binary search over the lanes' pre values with six shuffles.
constexpr int kWaveLogMax = 8;  // per-wave LDS hash: up to 256 buckets of four ids
constexpr int kWaveOwnerMax = 2 << kWaveLogMax;  // owner lists fill at most half the slots
constexpr int kWaveThreads = 256;  // four independent waves per workgroup

// Home bucket of w among 2^LOGB: a multiplicative hash, so ids sharing their
// low bits (strided or clustered labels) still spread.
template <int LOGB>
__device__ __forceinline__ uint32_t wave_bucket(int32_t w) {
  return ((uint32_t)w * 0x9E3779B1u) >> (32 - LOGB);
}

// Inserts key; returns whether a full bucket had to be passed (and was
// marked spilled, as in tab_insert).
template <int LOGB>
__device__ __forceinline__ bool wave_tab_insert(int4* tab, int32_t key) {
  uint32_t b = wave_bucket<LOGB>(key);
  bool passed = false;
  for (;;) {
    int* s = reinterpret_cast<int*>(tab + b);
    if (atomicCAS(s, kEmpty, key) == kEmpty) return passed;
    if (atomicCAS(s + 1, kEmpty, key) == kEmpty) return passed;
    if (atomicCAS(s + 2, kEmpty, key) == kEmpty) return passed;
    if (atomicCAS(s + 3, kEmpty, key) == kEmpty) return passed;
    atomicOr(s + 3, (int)0x80000000);
    passed = true;
    b = (b + 1) & ((1u << LOGB) - 1);
  }
}

// Continues a lookup past the spilled home bucket b; ends because at most half
// the slots are used, so some bucket is not full and therefore not spilled.
template <int LOGB>
__device__ __forceinline__ bool wave_tab_continue(const int4* tab, uint32_t b, int32_t w) {
  for (;;) {
    b = (b + 1) & ((1u << LOGB) - 1);
    const int4 t = tab[b];
    if (bucket_miss(t, w) == 0) return true;
    if (!bucket_spilled(t)) return false;
  }
}

// The same, returning the slot index 4*bucket + s of w, or -1.
template <int LOGB>
__device__ __forceinline__ int wave_tab_continue_slot(const int4* tab, uint32_t b, int32_t w) {
  for (;;) {
    b = (b + 1) & ((1u << LOGB) - 1);
    const int4 t = tab[b];
    const int sl = bucket_slot(t, b, w);
    if (sl >= 0) return sl;
    if (!bucket_spilled(t)) return -1;
  }
}

// Orders this wave's LDS accesses before and after it (one wave owns the
// table; LDS instructions of a wave execute in order).
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave tiles are tdesc[*t_begin .. *t_end).  Each wave draws runs of
// tile_run consecutive tiles from a 64-bit queue and keeps its table while
// the owner stays the same.  Loads are issued one round ahead: while round r
// of a tile is probed, the chunks of the next round -- of the same tile, or
// the first round of the next tile together with that tile's owner list --
// are in flight, and the items of the tile after and the descriptor after
// that are loaded earlier still.  Hits are counted per wave (ballot
// population counts) and added once per wave, or, in a batch, once per run of
// tiles of one graph.
struct WaveTile {
  int64_t i0, a0;  // first item; ooff[x]
  int32_t x, n, dx;
  bool valid;
};

// One lane's chunk of a round: four ids and how many of them lie inside the list.
struct WaveChunk {
  int4 v;
  int32_t r;
  int32_t y;  // the item's partner (owned-range and per-vertex credit)
  int32_t k;  // the item's index in the tile (per-vertex credit)
};

// CM: kCreditTotal; kCreditRange (owned-range counts: each hit {x, y, w} is
// split by its minimum and middle vertex, as in the group kernel); or
// kCreditVertex (per-vertex counts: each hit goes to min(x, y, w) -- to the
// table id w through a 32-bit counter per table slot, kept while the owner
// stays; to the partner y through a counter per item of the tile; to the
// owner x in registers -- all added to HBM when the owner or tile ends).
template <int CM, bool BATCH, int WPE, int LOGB>
__global__ __launch_bounds__(kWaveThreads) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_count_waves(
    const int32_t* __restrict__ ot, const Item* __restrict__ items, const TileDesc* __restrict__ tdesc,
    const int64_t* __restrict__ t_begin, const int64_t* __restrict__ t_end,
    uint64_t* __restrict__ counters, const int32_t* __restrict__ gid,
    unsigned long long* __restrict__ gtotal, uint32_t tile_run, int64_t rlo, int64_t rhi,
    unsigned long long* __restrict__ pv) {
  constexpr bool RANGE = CM == kCreditRange;
  constexpr bool PV = CM == kCreditVertex;
  constexpr int kWaves = kWaveThreads / kWave;
  constexpr int kB = 1 << LOGB;           // buckets of the wave's table
  constexpr int kOwnIds = 2 * kB / kWave;  // owner ids per lane (owner lists of at most 2 kB ids)
  static_assert(LOGB >= 6 && LOGB <= kWaveLogMax, "64 to 256 buckets");
  __shared__ int4 s_tab[kWaves][kB];
  __shared__ uint32_t s_wc[PV ? kWaves : 1][PV ? 4 * kB : 1];  // per table slot
  __shared__ uint32_t s_yc[PV ? kWaves : 1][PV ? kWaveItems : 1];        // per item of the tile
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  int4* tab = s_tab[wv];
  uint32_t* wc = s_wc[PV ? wv : 0];
  uint32_t* yc = s_yc[PV ? wv : 0];
  if constexpr (PV) {
#pragma unroll
    for (int i = 0; i < 4 * kB / kWave; i++) wc[lane + kWave * i] = 0;
    yc[lane] = 0;
  }
  const int64_t tb = *t_begin;
  const uint64_t nt = (uint64_t)(*t_end - tb);
  // Waves beyond the number of tiles have nothing to draw and leave at once
  // (each would otherwise add to all eight queue counters on its way out).
  if ((uint64_t)blockIdx.x * (kWaveThreads / kWave) + threadIdx.x / kWave >= nt) return;  // uniform per wave
  // One queue per XCD over an eighth of the tiles (consecutive tiles: the same
  // owners, and in a batch the same graphs, so they stay in that XCD's L2);
  // a wave whose XCD's share is drawn goes on to the next XCD's queue.
  unsigned long long* queues = reinterpret_cast<unsigned long long*>(&counters[kWaveQueue]);
  int q = (int)(blockIdx.x % kXcds), qleft = kXcds;
  uint64_t run_next = 0, run_end = 0;
  auto draw = [&]() -> uint64_t {
    while (run_next >= run_end) {
      if (qleft == 0) return nt;  // every queue drawn: no tile
      const uint64_t qlo = nt * (uint64_t)q / kXcds, qhi = nt * (uint64_t)(q + 1) / kXcds;
      unsigned long long v = 0;
      if (lane == 0) v = atomicAdd(&queues[q], (unsigned long long)tile_run);
      const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
      const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
      const uint64_t got = ((uint64_t)hi << 32) | lo;
      if (qlo + got < qhi) {
        run_next = qlo + got;
        run_end = min(run_next + tile_run, qhi);
      } else {
        q = (q + 1) % kXcds;
        qleft--;
      }
    }
    return run_next++;
  };
  const WaveTile none{0, 0, 0, 0, 0, false};
  const int4* __restrict__ td4 = reinterpret_cast<const int4*>(tdesc);
  auto load_desc = [&](uint64_t t) -> WaveTile {
    WaveTile d = none;
    if (t >= nt) return d;
    const int4 u0 = td4[2 * (tb + (int64_t)t)], u1 = td4[2 * (tb + (int64_t)t) + 1];
    auto uni64 = [](int32_t lo32, int32_t hi32) -> int64_t {
      return (int64_t)(((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)hi32) << 32) |
                       __builtin_amdgcn_readfirstlane((uint32_t)lo32));
    };
    d.i0 = uni64(u0.x, u0.y);
    d.a0 = uni64(u0.z, u0.w);
    d.x = __builtin_amdgcn_readfirstlane(u1.x);
    d.n = __builtin_amdgcn_readfirstlane(u1.y);
    d.dx = __builtin_amdgcn_readfirstlane(u1.z);
    d.valid = true;
    return d;
  };
  // Lane k holds item k of a tile: its list's offset and length (and partner).
  auto load_items = [&](const WaveTile& d, int64_t& ib, int32_t& il, int32_t& iy) {
    ib = 0;
    il = 0;
    iy = 0;
    if (d.valid && lane < d.n) {
      const Item it = items[d.i0 + lane];
      ib = it.base;
      il = it.len;
      iy = it.y;
    }
  };
  // Where each item's chunks start in the tile's run of chunks (lane k: item k).
  auto chunk_starts = [&](int32_t il, int32_t& pre, int32_t& total) {
    const int32_t nc = (il + 3) >> 2;
    int32_t inc = nc;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int32_t y = __shfl_up(inc, o, kWave);
      if (lane >= o) inc += y;
    }
    pre = inc - nc;
    total = __builtin_amdgcn_readlane(inc, kWave - 1);
  };
  // Chunk c0 + lane of a tile: its item kk is the last with pre[kk] <= c,
  // found by a binary search over the lanes' pre values (lane k holds pre[k],
  // ascending; items 0..n-1), six shuffles.
  auto load_round = [&](int64_t ib, int32_t il, int32_t iy, int32_t pre, int n, int32_t total,
                        int32_t c0) -> WaveChunk {
    const int32_t c = c0 + lane;
    int kk = 0;
    int32_t kp = 0;  // pre[0]
#pragma unroll
    for (int step = kWave / 2; step > 0; step >>= 1) {
      const int cand = kk + step;  // <= 63
      const int32_t pc = __shfl(pre, cand, kWave);
      if (cand < n && pc <= c) {
        kk = cand;
        kp = pc;
      }
    }
    const int64_t kb = __shfl(ib, kk, kWave);
    const int32_t kl = __shfl(il, kk, kWave);
    const int32_t q = c - kp;  // chunk of item kk
    WaveChunk ch{make_int4(-1, -1, -1, -1), kl - 4 * q, 0, kk};  // ids inside the list: <= 0 past the run's end
    if constexpr (RANGE || PV) ch.y = __shfl(iy, kk, kWave);
    if (c < total) {
      const v4i_a4 tq = *reinterpret_cast<const gv4i_a4*>(reinterpret_cast<uint64_t>(ot + kb + 4 * (int64_t)q));
      ch.v = make_int4(tq.x, tq.y, tq.z, tq.w);
    }
    return ch;
  };
  // An owner list of at most 2 kB ids: kOwnIds ids per lane.
  auto load_owner = [&](const WaveTile& d, int32_t (&o)[kOwnIds]) {
#pragma unroll
    for (int i = 0; i < kOwnIds; i++) o[i] = lane + kWave * i < d.dx ? ot[d.a0 + kWave * i + lane] : 0;
  };
  bool spl = false;  // some bucket of the table spilled (uniform)
  auto build_table = [&](const WaveTile& d, const int32_t (&o)[kOwnIds]) {
#pragma unroll
    for (int i = 0; i < kB / kWave; i++) tab[lane + kWave * i] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
    wave_lds_fence();
    bool passed = false;
#pragma unroll
    for (int i = 0; i < kOwnIds; i++)
      if (lane + kWave * i < d.dx) passed |= wave_tab_insert<LOGB>(tab, o[i]);
    spl = __builtin_amdgcn_ballot_w64(passed) != 0;
    wave_lds_fence();
  };
  uint64_t hits = 0;   // uniform: sums of ballot population counts
  uint64_t owned = 0, local = 0;  // RANGE, uniform: hits whose minimum / minimum and middle are in [rlo, rhi)
  int32_t cur_x = 0;   // owner of the tile being probed
  // Owned-range split of the hits in mask `hit` (ids w, partner y, owner cur_x).
  auto range_split = [&](uint64_t hit, int32_t w, int32_t y) {
    const int32_t ya = min(cur_x, y), yb = max(cur_x, y);
    const int32_t m = min(ya, w), mid = max(ya, min(yb, w));
    const uint64_t own = hit & __builtin_amdgcn_ballot_w64((uint32_t)(m - (int32_t)rlo) < (uint32_t)(rhi - rlo));
    owned += (uint64_t)__popcll(own);
    local += (uint64_t)__popcll(own & __builtin_amdgcn_ballot_w64(mid < (int32_t)rhi));
  };
  uint32_t at_x = 0;  // PV: hits credited to the owner (per lane, summed when the owner ends)
  // PV: one hit of w at table slot `slot`, for an item with partner y and index k.
  auto pv_hit = [&](int32_t w, int slot, int32_t y, uint32_t& ny) {
    if (w < min(cur_x, y))
      atomicAdd(&wc[slot], 1u);
    else if (cur_x < y)
      at_x++;
    else
      ny++;
  };
  auto probe_pv = [&](const WaveChunk& ch) {
    const int32_t w[4] = {ch.v.x, ch.v.y, ch.v.z, ch.v.w};
    int4 tt[4];
#pragma unroll
    for (int j = 0; j < 4; j++) tt[j] = tab[wave_bucket<LOGB>(w[j])];
    uint32_t ny = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool e0 = tt[j].x == w[j], e1 = tt[j].y == w[j], e2 = tt[j].z == w[j], e3 = tt[j].w == w[j];
      const bool h = j < ch.r && (e0 | e1 | e2 | e3);
      hits += (uint64_t)__popcll(__builtin_amdgcn_ballot_w64(h));
      if (h) pv_hit(w[j], (int)(wave_bucket<LOGB>(w[j]) << 2) + ((e1 | e3) ? 1 : 0) + ((e2 | e3) ? 2 : 0), ch.y, ny);
    }
    if (spl && __builtin_amdgcn_ballot_w64((tt[0].w | tt[1].w | tt[2].w | tt[3].w) < 0)) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        int slot = -1;
        if (j < ch.r && tt[j].w < 0) {
          if ((tt[j].w & 0x7fffffff) == w[j])
            slot = (int)(wave_bucket<LOGB>(w[j]) << 2) + 3;  // slot 3 of a spilled bucket: the raw compare missed it
          else if (bucket_miss(tt[j], w[j]) != 0)
            slot = wave_tab_continue_slot<LOGB>(tab, wave_bucket<LOGB>(w[j]), w[j]);
        }
        hits += (uint64_t)__popcll(__builtin_amdgcn_ballot_w64(slot >= 0));
        if (slot >= 0) pv_hit(w[j], slot, ch.y, ny);
      }
    }
    if (ny) atomicAdd(&yc[ch.k], ny);
  };
  // PV: partner credits of the tile just probed (lane k: item k, partner iy).
  auto flush_items = [&](int n, int32_t iy) {
    if constexpr (PV) {
      wave_lds_fence();
      if (lane < n) {
        const uint32_t c = yc[lane];
        if (c) {
          atomicAdd(&pv[iy], (unsigned long long)c);
          yc[lane] = 0;
        }
      }
      wave_lds_fence();
    }
  };
  // PV: the table's slot credits and the owner's credits, before the table goes.
  auto flush_owner = [&]() {
    if constexpr (PV) {
      wave_lds_fence();
#pragma unroll
      for (int i = 0; i < 4 * kB / kWave; i++) {
        const int sl = lane + kWave * i;
        const uint32_t c = wc[sl];
        if (c) {
          const int4 t = tab[sl >> 2];
          const int32_t key = (sl & 3) == 0 ? t.x : (sl & 3) == 1 ? t.y : (sl & 3) == 2 ? t.z : (t.w & 0x7fffffff);
          atomicAdd(&pv[key], (unsigned long long)c);
          wc[sl] = 0;
        }
      }
      const uint64_t ax = wave_sum_u64((uint64_t)at_x);  // lane 0 holds the sum
      if (lane == 0 && ax) atomicAdd(&pv[cur_x], (unsigned long long)ax);
      at_x = 0;
      wave_lds_fence();
    }
  };
  auto probe = [&](const WaveChunk& ch) {
    if constexpr (PV) {
      probe_pv(ch);
      return;
    }
    const int32_t w[4] = {ch.v.x, ch.v.y, ch.v.z, ch.v.w};
    int4 tt[4];
#pragma unroll
    for (int j = 0; j < 4; j++) tt[j] = tab[wave_bucket<LOGB>(w[j])];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint64_t hit = (__builtin_amdgcn_ballot_w64(tt[j].x == w[j]) | __builtin_amdgcn_ballot_w64(tt[j].y == w[j]) |
                            __builtin_amdgcn_ballot_w64(tt[j].z == w[j]) | __builtin_amdgcn_ballot_w64(tt[j].w == w[j])) &
                           __builtin_amdgcn_ballot_w64(j < ch.r);
      hits += (uint64_t)__popcll(hit);
      if constexpr (RANGE) range_split(hit, w[j], ch.y);
    }
    // A key in slot 3 of a spilled bucket carries the mark (the raw compare
    // missed it), and keys that passed a full bucket sit further on.
    if (spl && __builtin_amdgcn_ballot_w64((tt[0].w | tt[1].w | tt[2].w | tt[3].w) < 0)) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        bool h = false;
        if (j < ch.r && tt[j].w < 0) {
          h = (tt[j].w & 0x7fffffff) == w[j];
          if (!h && bucket_miss(tt[j], w[j]) != 0) h = wave_tab_continue<LOGB>(tab, wave_bucket<LOGB>(w[j]), w[j]);
        }
        const uint64_t hm = __builtin_amdgcn_ballot_w64(h);
        hits += (uint64_t)__popcll(hm);
        if constexpr (RANGE) range_split(hm, w[j], ch.y);
      }
    }
  };
  int32_t g_cur = -1;  // BATCH: graph of `hits`
  auto enter_graph = [&](const WaveTile& d) {
    if constexpr (BATCH) {
      const int32_t g = gid[d.x];
      if (g != g_cur) {
        if (lane == 0 && hits) atomicAdd(&gtotal[g_cur], (unsigned long long)hits);
        hits = 0;
        g_cur = g;
      }
    }
  };

  // Prologue: the first tile's owner list, items and first round; the next
  // tile's items; the descriptor after it.
  WaveTile dc = load_desc(draw());
  if (dc.valid) {
    int64_t ib_c, ib_n;
    int32_t il_c, il_n, iy_c, iy_n, own[kOwnIds];
    load_items(dc, ib_c, il_c, iy_c);
    load_owner(dc, own);
    WaveTile dn = load_desc(draw());
    load_items(dn, ib_n, il_n, iy_n);
    WaveTile dnn = dn.valid ? load_desc(draw()) : none;
    enter_graph(dc);
    build_table(dc, own);
    int64_t table_a0 = dc.a0;
    cur_x = dc.x;
    int32_t pre_c, total_c;
    chunk_starts(il_c, pre_c, total_c);
    WaveChunk cur = load_round(ib_c, il_c, iy_c, pre_c, dc.n, total_c, 0);
    int32_t c0 = 0;
    for (;;) {
      const bool same = c0 + kWave < total_c;  // uniform: the tile has another round
      WaveChunk nxt{make_int4(-1, -1, -1, -1), 0, 0, 0};
      int32_t pre_n = 0, total_n = 0;
      const bool fresh = !same && dn.valid && dn.a0 != table_a0;  // the next tile needs a new table
      if (same) {
        nxt = load_round(ib_c, il_c, iy_c, pre_c, dc.n, total_c, c0 + kWave);
      } else if (dn.valid) {
        chunk_starts(il_n, pre_n, total_n);
        nxt = load_round(ib_n, il_n, iy_n, pre_n, dn.n, total_n, 0);
        if (fresh) load_owner(dn, own);
      }
      probe(cur);
      cur = nxt;
      if (same) {
        c0 += kWave;
        continue;
      }
      flush_items(dc.n, iy_c);
      if (!dn.valid) break;  // uniform
      // Tile boundary: dn becomes the current tile.
      enter_graph(dn);
      if (fresh) {
        flush_owner();
        build_table(dn, own);
        table_a0 = dn.a0;
      }
      dc = dn;
      cur_x = dc.x;
      ib_c = ib_n;
      il_c = il_n;
      iy_c = iy_n;
      pre_c = pre_n;
      total_c = total_n;
      c0 = 0;
      dn = dnn;
      load_items(dn, ib_n, il_n, iy_n);
      dnn = dn.valid ? load_desc(draw()) : none;
    }
    flush_owner();
  }
  if (lane == 0 && hits) {
    if constexpr (BATCH)
      atomicAdd(&gtotal[g_cur], (unsigned long long)hits);
    else
      atomicAdd((unsigned long long*)&counters[kTotal], (unsigned long long)hits);
  }
  if constexpr (RANGE) {
    if (lane == 0 && owned) atomicAdd((unsigned long long*)&counters[kOwned], (unsigned long long)owned);
    if (lane == 0 && local) atomicAdd((unsigned long long*)&counters[kLocal], (unsigned long long)local);
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

// Tiles of all owners: at most one partial tile per owner plus the full ones,
// and never more than the heavy edges.
static int64_t tiles_bound(int64_t nv, int64_t ne) {
  const int64_t b = nv + ne / kMinItemsPerTile + 1;
  return b < ne + 1 ? b : ne + 1;
}

// Largest owner out-list counted in wave tiles, 0 for none.
// By default wave tiles (and, in batches, the low light cut that sends most
// edges to them) are used from kWaveMinEntries (one graph) or
// kBatchWaveMinEntries (a batch) entries on: below that the extra launch and
// the longer tile scan cost more than the small tiles do.  A value set with
// xt_ctx_tune / xt_ctx_tune_waves applies at every size.
static bool wave_size(int64_t ne, bool batch) {
  return ne >= (batch ? kBatchWaveMinEntries : kWaveMinEntries);
}
static int wave_owner_cut(const xt_ctx* ctx, bool batch, int64_t ne) {
  const int w = ctx->wave_owner != -1 ? ctx->wave_owner
                : !wave_size(ne, batch) ? 0
                : (batch ? kBatchWaveOwner : kWaveOwner);
  return w < 0 ? 0 : (w > kWaveOwnerMax ? kWaveOwnerMax : w);
}

template <typename T>
int count_oriented(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, const int64_t* in_off,
                   const T* in_tails, uint64_t* pv, const int32_t* gid, uint64_t* gtotal,
                   hipStream_t st, hipEvent_t* evs, const OwnedRange* range) {
  if (nv <= 0 || ne_bound <= 0) return XT_OK;
  // A tile's chunk indices are 32-bit: its lists hold at most ne_bound ids.
  if (ne_bound >= kMaxCountEntries) {
    return set_err(ctx, XT_ERR_CAPACITY,
                   "graph too large: the counting kernels take fewer than 2^33 - 2^16 entries");
  }
  const bool batch = gid != nullptr;
  if (range && (pv || batch))
    return set_err(ctx, XT_ERR_ARG, "an owned-range count takes neither per-vertex output nor a batch");
  const int64_t rlo = range ? range->lo : 0, rhi = range ? range->hi : 0;
  const int64_t nchunks = (ne_bound + kChunk - 1) / kChunk;
  XT_TRY(ensure(ctx, s.itcnt, (size_t)nv * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.items, (size_t)ne_bound * sizeof(Item)));
  XT_TRY(ensure(ctx, s.tiledesc, (size_t)tiles_bound(nv, ne_bound) * sizeof(TileDesc)));
  // Owners with out-lists of at most wown ids go to wave tiles (plain counts).
  const int wown = wave_owner_cut(ctx, batch, ne_bound);
  ctx->last_wave_owner = wown;
  XT_TRY(ensure(ctx, s.deg, (size_t)(wown ? 2 : 1) * (size_t)nv * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.tileoff, (size_t)((wown ? 2 : 1) * nv + 1) * sizeof(int64_t)));
  uint64_t* counters = s.counters.as<uint64_t>();
  XT_TRY(fill_u64(ctx, counters + kTileQueue, 1, 0, st));
  if (wown) XT_TRY(fill_u64(ctx, counters + kWaveQueue, kXcds, 0, st));
  XT_TRY(fill_u32(ctx, s.itcnt.as<uint32_t>(), nv, 0u, st));
  const int64_t* ooff = s.ooff.as<int64_t>();
  const int32_t* ot = s.otails.as<int32_t>();
  const uint8_t* oflags = s.oflags.as<uint8_t>();
  unsigned long long* pvp = (unsigned long long*)pv;
  unsigned long long* gt = (unsigned long long*)gtotal;
  const int cut = ctx->small_cut > 0 ? ctx->small_cut
                  : !gid                   ? kLightCut
                  : wave_size(ne_bound, true) ? kBatchLightCut
                                              : kBatchLightCutSmall;
  int64_t g = (nv + 1 + 255) / 256;
  if (g > 65536) g = 65536;
  // every oriented offset and out-degree fits 32 bits (unless the wide layout is forced)
  const bool narrow = ne_bound < (1ll << 32) && !ctx->force_wide;
  uint64_t* vrec = nullptr;
  uint4* vrange = nullptr;
  uint16_t* dp16 = nullptr;
  if (narrow && range) {
    XT_TRY(ensure(ctx, s.vrec, (size_t)nv * sizeof(uint4)));
    vrange = s.vrec.as<uint4>();
    hipLaunchKernelGGL(k_vertex_ranges, dim3((unsigned)g), dim3(256), 0, st, nv, ooff, ot, rlo, rhi, vrange);
  } else if (narrow) {
    XT_TRY(ensure(ctx, s.vrec, (size_t)nv * sizeof(uint64_t)));
    vrec = s.vrec.as<uint64_t>();
    hipLaunchKernelGGL(k_vertex_records, dim3((unsigned)g), dim3(256), 0, st, nv, ooff, vrec);
  } else {
    // s.deg (the orientation's degree keys until now) takes the 16-bit
    // out-degrees, later the tile counts.
    dp16 = s.deg.as<uint16_t>();
    hipLaunchKernelGGL(k_dplus16, dim3((unsigned)g), dim3(256), 0, st, nv, ooff, dp16);
  }
  XT_HIP(ctx, hipGetLastError());
#define XT_LIGHT(N, C, B)                                                                           \
  hipLaunchKernelGGL((k_light_items<T, N, C, B>), dim3(xcd_grid(nchunks)), dim3(kBlock), 0, st, nv,  \
                     ne_bound, in_off, in_tails, oflags, ooff, (const uint64_t*)vrec,              \
                     (const uint4*)vrange, (const uint16_t*)dp16, ot, (const int64_t*)s.crows.as<int64_t>(), cut, \
                     counters, s.itcnt.as<uint32_t>(), s.items.as<Item>(), pvp, gid, gt, rlo, rhi)
#define XT_LIGHT_PB(N)                    \
  if (range)                              \
    XT_LIGHT(N, kCreditRange, false);     \
  else if (pv && batch)                   \
    XT_LIGHT(N, kCreditVertex, true);     \
  else if (pv)                            \
    XT_LIGHT(N, kCreditVertex, false);    \
  else if (batch)                         \
    XT_LIGHT(N, kCreditTotal, true);      \
  else                                    \
    XT_LIGHT(N, kCreditTotal, false);
  if (narrow) {
    XT_LIGHT_PB(true);
  } else {
    XT_LIGHT_PB(false);
  }
#undef XT_LIGHT_PB
#undef XT_LIGHT
  XT_HIP(ctx, hipGetLastError());
  if (evs) XT_HIP(ctx, hipEventRecord(evs[0], st));
  const int ipt = ctx->items_per_tile;
  hipLaunchKernelGGL(k_tile_counts, dim3((unsigned)g), dim3(256), 0, st, nv,
                     (const uint32_t*)s.itcnt.as<uint32_t>(), ooff, s.deg.as<uint32_t>(), ipt, wown);
  XT_HIP(ctx, hipGetLastError());
  XT_TRY(scan_u32_to_i64(ctx, s, s.deg.as<uint32_t>(), (wown ? 2 : 1) * nv, s.tileoff.as<int64_t>(), st));
  hipLaunchKernelGGL(k_tile_desc, dim3((unsigned)g), dim3(256), 0, st, nv, in_off, ooff,
                     (const uint32_t*)s.itcnt.as<uint32_t>(), (const int64_t*)s.tileoff.as<int64_t>(),
                     s.tiledesc.as<TileDesc>(), ipt, wown);
  XT_HIP(ctx, hipGetLastError());
  if (evs) XT_HIP(ctx, hipEventRecord(evs[1], st));
  const unsigned groups = (unsigned)(ctx->num_cus * ctx->group_blocks);
#define XT_GROUPS_H(C, B, R, W, M)                                                                 \
  hipLaunchKernelGGL((k_count_groups<C, B, R, W, M>), dim3(groups), dim3(kGroupThreads), 0, st, ot,  \
                     (const Item*)s.items.as<Item>(), (const TileDesc*)s.tiledesc.as<TileDesc>(),   \
                     (const int64_t*)s.tileoff.as<int64_t>() + nv, counters, pvp, gid, gt, rlo, rhi,   \
                     (uint32_t)ctx->tile_run)
#define XT_GROUPS(C, B, R, W)                \
  do {                                       \
    if (ctx->group_hash == kHashLowBits)     \
      XT_GROUPS_H(C, B, R, W, false);        \
    else                                     \
      XT_GROUPS_H(C, B, R, W, true);         \
  } while (0)
#define XT_GROUPS_R(P, B)          \
  if (ctx->group_round == 1)       \
    XT_GROUPS(P, B, 1, 8);         \
  else if (ctx->group_round == 2)  \
    XT_GROUPS(P, B, 2, 6);         \
  else                             \
    XT_GROUPS(P, B, 4, 4);
  if (range) {
    XT_GROUPS_R(kCreditRange, false);
  } else if (pv && batch) {
    XT_GROUPS_R(kCreditVertex, true);
  } else if (pv) {
    XT_GROUPS_R(kCreditVertex, false);
  } else if (batch) {
    XT_GROUPS_R(kCreditTotal, true);
  } else {
    XT_GROUPS_R(kCreditTotal, false);
  }
#undef XT_GROUPS_R
#undef XT_GROUPS
#undef XT_GROUPS_H
  XT_HIP(ctx, hipGetLastError());
  if (wown) {
    const int64_t* tof = s.tileoff.as<int64_t>();
    const unsigned wblocks = (unsigned)(ctx->num_cus * kWaveBlocksPerCU);
#define XT_WAVES_L(C, B, L)                                                                               \
  hipLaunchKernelGGL((k_count_waves<C, B, C == kCreditVertex ? 5 : C == kCreditRange ? 7 : 8, L>), dim3(wblocks),    \
                     dim3(kWaveThreads), 0, st, ot,                                                             \
                     (const Item*)s.items.as<Item>(), (const TileDesc*)s.tiledesc.as<TileDesc>(), tof + nv,     \
                     tof + 2 * nv, counters, gid, gt, (uint32_t)ctx->tile_run, rlo, rhi, pvp)
#define XT_WAVES(C, B)            \
  do {                            \
    if (wown <= 128)              \
      XT_WAVES_L(C, B, 6);        \
    else if (wown <= 256)         \
      XT_WAVES_L(C, B, 7);        \
    else                          \
      XT_WAVES_L(C, B, 8);        \
  } while (0)
    if (range)
      XT_WAVES(kCreditRange, false);
    else if (pv && batch)
      XT_WAVES(kCreditVertex, true);
    else if (pv)
      XT_WAVES(kCreditVertex, false);
    else if (batch)
      XT_WAVES(kCreditTotal, true);
    else
      XT_WAVES(kCreditTotal, false);
#undef XT_WAVES
#undef XT_WAVES_L
    XT_HIP(ctx, hipGetLastError());
  }
  return XT_OK;
}

template int count_oriented<int32_t>(xt_ctx*, Scratch&, int64_t, int64_t, const int64_t*, const int32_t*,
                                     uint64_t*, const int32_t*, uint64_t*, hipStream_t, hipEvent_t*,
                                     const OwnedRange*);
template int count_oriented<int64_t>(xt_ctx*, Scratch&, int64_t, int64_t, const int64_t*, const int64_t*,
                                     uint64_t*, const int32_t*, uint64_t*, hipStream_t, hipEvent_t*,
                                     const OwnedRange*);

int alg_stats(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, hipStream_t st, int cut) {
  if (nv <= 0 || ne_bound <= 0) return XT_OK;
  const int64_t nchunks = (ne_bound + kChunk - 1) / kChunk;
  uint64_t* counters = s.counters.as<uint64_t>();
  XT_TRY(fill_u64(ctx, counters + kStatMin, 1, 0, st));
  XT_TRY(fill_u64(ctx, counters + kStatMinHeavy, 2, 0, st));
  hipLaunchKernelGGL(k_alg_stats, dim3((unsigned)nchunks), dim3(kBlock), 0, st, nv,
                     (const int64_t*)s.ooff.as<int64_t>(), (const int32_t*)s.otails.as<int32_t>(),
                     cut > 0 ? cut : ctx->small_cut > 0 ? ctx->small_cut : kLightCut, counters);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

// Every scratch buffer a count of a graph of up to (nv, ne) entries can
// request, with the largest size orient(), count_oriented() and the scans ask
// for: f(buffer, bytes).
template <typename F>
static void for_each_need(Scratch& s, int64_t nv, int64_t ne, bool per_vertex, F&& f) {
  const int64_t nvv = nv > 0 ? nv : 1, nee = ne > 0 ? ne : 1;
  const int64_t nchunks = (nee + kChunk - 1) / kChunk;
  f(s.deg, (size_t)(2 * nvv) * sizeof(uint32_t));  // degree keys, then tile counts (group and wave)
  f(s.dplus, (size_t)nvv * sizeof(uint32_t));
  f(s.tailcnt, (size_t)nchunks * sizeof(uint32_t));
  f(s.qscan, (size_t)(nchunks + 1) * sizeof(int64_t));
  f(s.ooff, (size_t)(nvv + 1) * sizeof(int64_t));
  f(s.otails, (size_t)(nee + kOtailsPad) * sizeof(int32_t));
  f(s.counters, kNumSlots * sizeof(uint64_t));
  f(s.oflags, (size_t)(nchunks * kBlock));
  f(s.crows, (size_t)(2 * nchunks) * sizeof(int64_t));
  f(s.itcnt, (size_t)nvv * sizeof(uint32_t));
  f(s.items, (size_t)nee * sizeof(Item));
  f(s.tileoff, (size_t)(2 * nvv + 1) * sizeof(int64_t));
  f(s.tiledesc, (size_t)tiles_bound(nvv, nee) * sizeof(TileDesc));
  if (nee < (1ll << 32)) f(s.vrec, (size_t)nvv * sizeof(uint64_t));
  f(s.scan_tmp, scan_tmp_bytes(2 * nvv) > scan_tmp_bytes(nchunks) ? scan_tmp_bytes(2 * nvv) : scan_tmp_bytes(nchunks));
  if (per_vertex) f(s.pv, (size_t)nvv * sizeof(uint64_t));
}

int reserve(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne, bool per_vertex) {
  int rc = XT_OK;
  for_each_need(s, nv, ne, per_vertex, [&](DevBuf& b, size_t bytes) {
    if (rc == XT_OK) rc = ensure(ctx, b, bytes);
  });
  return rc;
}

static size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t carve_bytes(int64_t nv, int64_t ne) {
  Scratch tmp;
  size_t tot = 0;
  for_each_need(tmp, nv, ne, false, [&](DevBuf&, size_t bytes) { tot += round256(bytes); });
  return tot;
}

void carve(Scratch& s, unsigned char* base, int64_t nv, int64_t ne) {
  size_t at = 0;
  for_each_need(s, nv, ne, false, [&](DevBuf& b, size_t bytes) {
    b.p = base + at;
    b.bytes = round256(bytes);
    b.fixed = true;
    at += b.bytes;
  });
}

void release(Scratch& s) {
  DevBuf* bufs[] = {&s.deg, &s.dplus, &s.tailcnt, &s.qscan, &s.ooff, &s.otails, &s.counters,
                    &s.oflags, &s.crows, &s.itcnt, &s.items, &s.tileoff, &s.tiledesc, &s.vrec, &s.pv, &s.scan_tmp, &s.in_off,
                    &s.in_tails};
  for (DevBuf* b : bufs) {
    if (b->fixed) continue;
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

int launch_add_total(xt_ctx* ctx, int64_t* out, hipStream_t st, const uint64_t* counters) {
  if (!counters) counters = ctx->s.counters.as<uint64_t>();
  hipLaunchKernelGGL(k_add_total, dim3(1), dim3(64), 0, st, counters, (unsigned long long*)out);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

}  // namespace xt
