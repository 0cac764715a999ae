// xt_internal.hpp -- shared internals of libxtric (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/xtric.h"
#include "xt_guard.hpp"

namespace xt {

// Launch geometry shared by the chunked kernels: one 256-thread workgroup
// (four wave64s) owns a contiguous chunk of kChunk CSR entries, eight per
// thread.  kOffCap bounds the vertex span staged in LDS per chunk.
constexpr int kBlock = 256;
constexpr int kPerThread = 8;
constexpr int kChunk = kBlock * kPerThread;  // 2048
// The oriented ids are read in 16-byte chunks from each list's first id; the
// last chunk of the last list may read up to three ids past |E+|.
constexpr int64_t kOtailsPad = 4;
constexpr int kLightCut = 32;       // default light cut, one graph
constexpr int kBatchLightCut = 8;   // default light cut, batches and the server (with wave tiles)
constexpr int kBatchLightCutSmall = 64;  // ... below kBatchWaveMinEntries (no wave tiles)
constexpr int kOffCap = 2048;
// Wave tiles (k_count_waves): owners whose out-list has at most this many ids
// are counted one wave per tile; 0 turns them off.
constexpr int kHashMix = 1;      // group-kernel table hash: top bits of id * 0x9E3779B1 (default)
constexpr int kHashLowBits = 2;  // ... or id mod 2048
constexpr int kWaveOwner = 256;      // default, one graph
constexpr int kBatchWaveOwner = 64;  // default, batches and the server
constexpr int kWaveBlocksPerCU = 8;  // persistent 256-thread workgroups per CU
// Default sizes (directed entries) from which wave tiles are used.
constexpr int64_t kWaveMinEntries = int64_t(1) << 27;       // one graph (R-MAT s23 on: 27.9 -> 26.2 ms)
constexpr int64_t kBatchWaveMinEntries = int64_t(1) << 22;  // a batch

// Counter slots in Scratch::counters (uint64).
enum Slot : int {
  kTotal = 0,          // global triangle count
  kBadIds = 3,         // neighbour ids outside [0, nv)
  kBadRows = 4,        // rows not strictly ascending
  kAsym = 5,           // entries without their reverse entry
  kBadOff = 6,         // offsets not monotone / wrong ends
  kStatMin = 7,        // sum of min(d+(u), d+(v)) over oriented edges (statistics kernel)
  kTileQueue = 8,      // next tile of the persistent group kernel
  kStatMinHeavy = 9,   // same sum over heavy edges only
  kStatHeavyEdges = 10,// number of heavy edges
  kHeavyItems = 11,    // heavy edges listed under their owners by the counting pass
  kOwned = 12,         // owned-range count: triangles whose minimum vertex is owned
  kLocal = 13,         // ... and whose middle vertex is owned too
  kWaveQueue = 14,     // next wave tile of k_count_waves, one queue per XCD (14..21)
  kNumSlots = 24
};

// Orientation rules.
enum Orient : int { kOrientDegree = 0, kOrientId = 1 };

// A grow-only device buffer owned by a context.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool fixed = false;  // a slice of caller-provided memory: never allocated, grown or freed
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Phase events recorded on the context stream.
enum Ev : int { kEvStart = 0, kEvOriented, kEvLight, kEvItems, kEvCounted, kNumEv };

struct Scratch {
  DevBuf deg;       // uint32[2nv]         degree of each vertex; later tiles per owner (group, wave)
  DevBuf dplus;     // uint32[nv]          out-degree of each vertex
  DevBuf tailcnt;   // uint32[nchunks]     flagged entries of each chunk's last vertex
  DevBuf qscan;     // int64[nchunks+1]    exclusive scan of tailcnt
  DevBuf ooff;      // int64[nv+1]         oriented offsets
  DevBuf vrec;      // uint64[nv]          ooff[x] | d+(x) << 32 when |E+| < 2^32 (gathered per entry)
  DevBuf otails;    // int32[ne+4]         oriented out-lists, ascending ids (|E+| <= ne), padded
  DevBuf counters;  // uint64[16]          total, tile queue, input-error flags, statistics
  DevBuf oflags;    // uint8[ne/8]         kept-entry mask per 8 CSR entries (orientation)
  DevBuf crows;     // int64[2*nchunks]    first and last row of each kChunk-entry input chunk
  DevBuf itcnt;     // uint32[nv]          heavy edges owned by each vertex
  DevBuf items;     // 16 B[ne]            item lists: owner x's {ooff[y], d+(y), y} records from items[off[x]]
  DevBuf tileoff;   // int64[2nv+1]        exclusive scan of tiles per owner (group tiles, then wave tiles)
  DevBuf tiledesc;  // 32 B[nv+ne/256+1]  {first item, ooff[x], x, items, d+(x)} of each tile
  DevBuf pv;        // uint64[nv]          per-vertex counts
  DevBuf scan_tmp;  // int64[..]           block sums of the scans
  DevBuf in_off;    // int64[nv+1]         staged input offsets (xt_count_csr)
  DevBuf in_tails;  // int32[ne]           staged input ids (xt_count_csr)
};

}  // namespace xt

struct xt_ctx {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  std::string err;
  std::recursive_mutex mu;
  xt::Scratch s;
  hipEvent_t ev[xt::kNumEv] = {};
  xt_timing timing = {};
  // Light edges: both out-lists at most this long; 0 picks kLightCut for one
  // graph and kBatchLightCut for batches (where the rest goes to wave tiles).
  int small_cut = 0;
  int group_blocks = 4;   // persistent group-kernel workgroups per CU
  int group_round = 2;    // 16-byte chunks per thread per group-kernel streaming round (1, 2, 4)
  int group_hash = 1;     // group-kernel LDS hash: kHashMix (multiplicative) or kHashLowBits (id mod 2048)
  int items_per_tile = 256;  // items of one group tile (64, 128 or 256)
  int tile_run = 16;         // consecutive tiles a group-kernel workgroup draws at once (1..64)
  int wave_owner = -1;       // wave tiles for owner lists up to this long; -1 default, 0 off
  int last_wave_owner = 0;  // wave-owner cut of the last count (tiles [tileoff[nv], tileoff[2nv]) when > 0)
  bool force_wide = false;  // use the 64-bit offset layout of graphs with >= 2^32 entries everywhere
  int fail_allocs = 0;      // failure injection: scratch allocations left to fail
};

struct xt_graph {
  int64_t nv = 0;
  int64_t ne = 0;
  int64_t* off = nullptr;   // int64[nv+1]
  int32_t* tails = nullptr; // int32[ne]
};

namespace xt {

int set_err(xt_ctx* ctx, int code, const std::string& msg);
int hip_err(xt_ctx* ctx, hipError_t e, const char* what);
// Ensure a scratch buffer holds at least `bytes`; contents are not preserved.
int ensure(xt_ctx* ctx, DevBuf& b, size_t bytes);

// ---- pipeline pieces (kernels live in the .hip files) ----
// Exclusive scan of n uint32 -> int64 out[0..n] (out[n] = total).
int scan_u32_to_i64(xt_ctx* ctx, Scratch& s, const uint32_t* in, int64_t n, int64_t* out,
                    hipStream_t st);
// The same with a caller-provided block-sum buffer of scan_tmp_bytes(n) bytes.
int scan_u32_to_i64_tmp(xt_ctx* ctx, const uint32_t* in, int64_t n, int64_t* out, int64_t* tmp,
                        hipStream_t st);

// Orientation of a symmetric CSR into the context's ooff/otails.
// Upper bound of |E+| is ne/2 (one direction per undirected pair).
template <typename T>
int orient(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne, const int64_t* off, const T* tails, int mode,
           hipStream_t st);

// Intersection + reduction over the context's oriented CSR.  Adds the global
// count to ctx counters[0]; per-vertex counts into pv when pv != nullptr.
// Batch mode: gid (graph of each vertex) and gtotal (per-graph totals) non-null
// and the intersection kernel drains a device tile queue with a fixed grid.
// evs (nullable): two events recorded after the light-edge kernel and after
// the item lists are built.
// Owned vertex range [lo, hi) of a credit-by-range count.
struct OwnedRange {
  int64_t lo, hi;
};

template <typename T>
int count_oriented(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, const int64_t* in_off,
                   const T* in_tails, uint64_t* pv, const int32_t* gid, uint64_t* gtotal,
                   hipStream_t st, hipEvent_t* evs, const OwnedRange* range = nullptr);
// Sums of min(d+(u), d+(v)) over the oriented edges of s (all, and those whose
// longer list exceeds `cut`: 0 = the context's light cut for one graph).
int alg_stats(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne_bound, hipStream_t st, int cut = 0);
// Pre-sizes every scratch buffer for graphs up to (nv, ne).
int reserve(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne, bool per_vertex);
// Bytes a single-graph count of (nv, ne) needs when every scratch buffer is a
// 256-byte-aligned slice of one region (global count, device input CSR).
size_t carve_bytes(int64_t nv, int64_t ne);
// Points every buffer a global count of (nv, ne) uses at a slice of
// [base, base + carve_bytes(nv, ne)), marked fixed so that no later request
// can allocate.  The caller checks the region is large enough.
void carve(Scratch& s, unsigned char* base, int64_t nv, int64_t ne);
void release(Scratch& s);

size_t scan_tmp_bytes(int64_t n);

// Fill kernels (used instead of hipMemsetAsync so the sequence can be captured).
int fill_u32(xt_ctx* ctx, uint32_t* p, int64_t n, uint32_t v, hipStream_t st);
int fill_u64(xt_ctx* ctx, uint64_t* p, int64_t n, uint64_t v, hipStream_t st);

// Validation of a device CSR; returns XT_ERR_INPUT with a message on failure.
// counters (nullable): the uint64[kNumSlots] slots to use, else the context's.
template <typename T>
int validate(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* off, const T* tails,
             bool symmetry, hipStream_t st, uint64_t* counters = nullptr);

// Whole single-graph count on the device (orientation + intersection).
template <typename T>
int count_device_csr(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne, const int64_t* off, const T* tails,
                     uint32_t flags, hipStream_t st, bool timed, const OwnedRange* range = nullptr);

// Device generators.
int gen_rmat(xt_ctx* ctx, int scale, int edgefactor, const uint32_t thr[3], uint64_t seed,
             int permute, xt_graph* g);
int gen_er(xt_ctx* ctx, int64_t n, uint64_t seed, uint64_t thr64, xt_graph* g);
// Edge list (host pairs) -> symmetric, merged, self-loop-free CSR on the device.
int graph_from_edges(xt_ctx* ctx, int64_t nv, int64_t m, const int64_t* src, const int64_t* dst, xt_graph* g);
// A resident CSR rebuilt as a simple undirected graph in place: rows sorted,
// duplicates merged, self-loops dropped, missing reverse entries added.
int graph_clean(xt_ctx* ctx, xt_graph* g);
// Rows of [lo, hi) and of their neighbours kept, the others emptied (fills out).
int graph_slice(xt_ctx* ctx, const xt_graph* g, int64_t lo, int64_t hi, xt_graph* out);

// Host mixing function shared by the generators (SplitMix64 output function).
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace xt

#define XT_HIP(ctx, expr)                                   \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ::xt::hip_err(ctx, _e, #expr); \
  } while (0)

#define XT_TRY(expr)          \
  do {                        \
    int _rc = (expr);         \
    if (_rc != XT_OK) return _rc; \
  } while (0)
