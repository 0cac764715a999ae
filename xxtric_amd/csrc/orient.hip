// orient.hip -- device orientation of a symmetric CSR, scans and fills (gfx950).
//
// Orientation keeps each undirected edge once, pointing from the lower to the
// higher key: key(x) = (deg x, x) for XT_ORIENT_DEGREE, key(x) = x for
// XT_ORIENT_ID (the reference's rule, xxtric.hpp:165 / xxtric-gpu.hpp:84).
// Out-lists keep ascending ids because they are filtered from sorted rows.
//
// Work split: the CSR entries are cut into fixed chunks of kChunk (2048)
// entries, one 256-thread workgroup per chunk, eight contiguous entries per
// thread.  A hub row therefore spans many workgroups instead of serialising on
// one wave.  The rows a chunk touches are staged in LDS when they fit; the
// first and last row of every chunk come from a table built once per count
// (k_chunk_rows), not from a search per workgroup.
//
//  k_orient_count: out-degree of every row (LDS counters, one global atomic per
//                  row and chunk) and, per chunk, the out-entries of the chunk's
//                  last row (tailcnt) -- the carry needed by rows that continue
//                  into the next chunk.
//  scans:          ooff = excl_scan(dplus), q = excl_scan(tailcnt).
//  k_orient_fill:  a chunk's kept entries fill one contiguous range of the
//                  oriented ids, from ooff[u] + (kept entries of its first row u
//                  in earlier chunks = q[c] - q[first chunk of u]); each entry
//                  goes to that start + its rank in the chunk (a block scan).
#include "xt_device.hpp"
#include "xt_internal.hpp"

namespace xt {



// ---------------------------------------------------------------------------
// Degrees saturated at 65535: the key (min(deg, 65535), id) is still a strict
// total order (ties broken by id), so the orientation stays a DAG and the
// count is unchanged; 16-bit keys halve the footprint of the neighbour gathers.
__global__ __launch_bounds__(256) void k_degree(int64_t nv, const int64_t* __restrict__ off,
                                                uint16_t* __restrict__ deg) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nv; u += stride) {
    const int64_t d = off[u + 1] - off[u];
    deg[u] = (uint16_t)(d < 65535 ? d : 65535);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, int64_t n, T v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_check_offsets(int64_t nv, int64_t ne, const int64_t* __restrict__ off,
                                uint64_t* __restrict__ counters) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t bad = 0;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nv; u += stride)
    if (off[u] > off[u + 1]) bad++;
  if (blockIdx.x == 0 && threadIdx.x == 0 && (off[0] != 0 || off[nv] != ne)) bad++;
  if (bad) atomicAdd((unsigned long long*)&counters[kBadOff], (unsigned long long)bad);
}

template <int MODE>
__device__ __forceinline__ uint64_t orient_key(const uint16_t* __restrict__ deg, int64_t x) {
  if constexpr (MODE == kOrientDegree)
    return ((uint64_t)deg[x] << 32) | (uint64_t)x;
  else
    return (uint64_t)x;
}

// Stages off[ulo .. uhi+1] of the chunk's rows in LDS when they fit.
struct ChunkRows {
  int64_t ulo, uhi, span;
  bool in_lds;
};

// Rows of chunk c from the per-chunk row table, offsets staged in LDS.
__device__ __forceinline__ ChunkRows chunk_rows_at(const int64_t* __restrict__ off,
                                                   const int64_t* __restrict__ crows, int64_t nv,
                                                   int64_t c, int64_t* s_off) {
  const RowSpan sp = chunk_span(crows, c, nv);
  ChunkRows r;
  r.ulo = sp.lo;
  r.uhi = sp.hi;
  r.span = r.uhi - r.ulo + 1;
  r.in_lds = r.span <= kOffCap;
  if (r.in_lds)
    for (int64_t k = threadIdx.x; k <= r.span; k += blockDim.x) s_off[k] = off[r.ulo + k];
  __syncthreads();
  return r;
}

// The same found by binary search (validation runs before the table exists).
__device__ __forceinline__ ChunkRows chunk_rows(const int64_t* __restrict__ off, int64_t nv,
                                                int64_t eb, int64_t ee, int64_t* s_off,
                                                int64_t* s_lohi) {
  if (threadIdx.x == 0) {
    const int64_t lo = row_of(off, 0, nv - 1, eb);
    s_lohi[0] = lo;
    s_lohi[1] = row_of(off, lo, nv - 1, ee - 1);
  }
  __syncthreads();
  ChunkRows r;
  r.ulo = s_lohi[0];
  r.uhi = s_lohi[1];
  r.span = r.uhi - r.ulo + 1;
  r.in_lds = r.span <= kOffCap;
  if (r.in_lds)
    for (int64_t k = threadIdx.x; k <= r.span; k += blockDim.x) s_off[k] = off[r.ulo + k];
  __syncthreads();
  return r;
}

__device__ __forceinline__ int64_t chunk_row_of(const ChunkRows& r, const int64_t* s_off,
                                                const int64_t* __restrict__ off, int64_t e) {
  return r.in_lds ? r.ulo + lds_row_of(s_off, (int)r.span, e) : row_of(off, r.ulo, r.uhi, e);
}

__device__ __forceinline__ int64_t chunk_off(const ChunkRows& r, const int64_t* s_off,
                                             const int64_t* __restrict__ off, int64_t u) {
  return r.in_lds ? s_off[u - r.ulo] : off[u];
}

// First and last row of every kChunk-entry chunk: crows[2c] holds the row of
// entry c*kChunk, crows[2c+1] the row of the chunk's last entry.  One thread
// per row writes the chunk boundaries that fall inside the row, so every slot
// is written once when the index is valid (rows outside [0, ne] are skipped).
__global__ __launch_bounds__(256) void k_chunk_rows(int64_t nv, int64_t ne, const int64_t* __restrict__ off,
                                                    int64_t* __restrict__ crows) {
  const int64_t nchunks = (ne + kChunk - 1) / kChunk;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nv; u += stride) {
    const int64_t a = off[u], b = off[u + 1];
    if (a < 0 || b > ne || a >= b) continue;
    for (int64_t c = (a + kChunk - 1) / kChunk; c * kChunk < b; c++) crows[2 * c] = u;  // c*K in [a, b)
    for (int64_t c1 = a / kChunk + 1; c1 * kChunk <= b; c1++) crows[2 * c1 - 1] = u;   // (c1*K - 1) in [a, b)
    if (b == ne) crows[2 * nchunks - 1] = u;  // the last chunk ends at entry ne - 1
  }
}

// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_orient_count(
    int64_t nv, int64_t ne, const int64_t* __restrict__ off, const T* __restrict__ tails,
    const uint16_t* __restrict__ deg, const int64_t* __restrict__ crows, uint32_t* __restrict__ dplus,
    uint32_t* __restrict__ tailcnt, uint64_t* __restrict__ counters,
    uint8_t* __restrict__ oflags) {
  __shared__ int64_t s_off[kOffCap + 1];
  __shared__ uint32_t s_cnt[kOffCap];
  __shared__ uint32_t s_tail;
  const int64_t nchunks = (ne + kChunk - 1) / kChunk;
  const int64_t c = xcd_block(nchunks);
  if (c >= nchunks) return;  // whole workgroup
  const int64_t eb = c * kChunk;
  const int64_t ee = min(ne, eb + (int64_t)kChunk);
  if (threadIdx.x == 0) s_tail = 0;
  const ChunkRows r = chunk_rows_at(off, crows, nv, c, s_off);
  if (r.in_lds)
    for (int64_t k = threadIdx.x; k < r.span; k += kBlock) s_cnt[k] = 0;
  __syncthreads();

  const int64_t e0 = eb + (int64_t)threadIdx.x * kPerThread;
  uint32_t bad = 0;
  if (e0 < ee) {
    int64_t vals[kPerThread];
    load_run<T>(tails, e0, ee, vals);
    // All eight neighbour keys are gathered before any is used (eight loads in
    // flight instead of one); ids out of range read vertex 0 and are counted bad.
    uint16_t dv[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; j++) {
      const int64_t v = vals[j];
      const bool ok = e0 + j < ee && v >= 0 && v < nv;
      dv[j] = MODE == kOrientDegree ? deg[ok ? v : 0] : 0;
    }
    int64_t u = chunk_row_of(r, s_off, off, e0);
    int64_t cur = chunk_off(r, s_off, off, u);
    int64_t nxt = chunk_off(r, s_off, off, u + 1);
    // key of the row itself, from its offsets (same saturation as k_degree)
    auto row_key = [&](int64_t uu, int64_t d) -> uint64_t {
      if constexpr (MODE == kOrientDegree)
        return ((uint64_t)(d < 65535 ? d : 65535) << 32) | (uint64_t)uu;
      else
        return (uint64_t)uu;
    };
    uint64_t ku = row_key(u, nxt - cur);
    uint32_t run = 0, fl = 0;
    auto flush = [&](int64_t uu, uint32_t cnt) {
      if (cnt == 0) return;
      if (r.in_lds)
        atomicAdd(&s_cnt[uu - r.ulo], cnt);
      else {
        atomicAdd(&dplus[uu], cnt);
        if (uu == r.uhi) atomicAdd(&s_tail, cnt);
      }
    };
#pragma unroll
    for (int j = 0; j < kPerThread; j++) {
      const int64_t e = e0 + j;
      if (e < ee) {
        while (e >= nxt) {
          flush(u, run);
          run = 0;
          ++u;
          cur = nxt;
          nxt = chunk_off(r, s_off, off, u + 1);
          ku = row_key(u, nxt - cur);
        }
        const int64_t v = vals[j];
        const uint64_t kv = MODE == kOrientDegree ? (((uint64_t)dv[j] << 32) | (uint64_t)v) : (uint64_t)v;
        if (v < 0 || v >= nv) {
          bad++;
        } else if (kv > ku) {
          run++;
          fl |= 1u << j;
        }
      }
    }
    flush(u, run);
    oflags[e0 / kPerThread] = (uint8_t)fl;  // kept entries of this run, for the fill pass
  }
  if (bad) atomicAdd((unsigned long long*)&counters[kBadIds], (unsigned long long)bad);
  __syncthreads();
  if (r.in_lds) {
    for (int64_t k = threadIdx.x; k < r.span; k += kBlock) {
      const uint32_t cnt = s_cnt[k];
      if (cnt) atomicAdd(&dplus[r.ulo + k], cnt);
    }
    if (threadIdx.x == 0) tailcnt[c] = s_cnt[r.span - 1];
  } else if (threadIdx.x == 0) {
    tailcnt[c] = s_tail;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_orient_fill(
    int64_t nv, int64_t ne, const int64_t* __restrict__ off, const T* __restrict__ tails,
    const uint8_t* __restrict__ oflags, const int64_t* __restrict__ ooff,
    const int64_t* __restrict__ q, const int64_t* __restrict__ crows, int32_t* __restrict__ otails) {
  __shared__ int64_t s_scan[kBlock / kWave];
  const int64_t nchunks = (ne + kChunk - 1) / kChunk;
  const int64_t c = xcd_block(nchunks);
  if (c >= nchunks) return;  // whole workgroup
  const int64_t eb = c * kChunk;
  const int64_t ee = min(ne, eb + (int64_t)kChunk);
  const int64_t e0 = eb + (int64_t)threadIdx.x * kPerThread;
  int64_t vals[kPerThread];
  uint32_t fl = 0;
  if (e0 < ee) {
    fl = oflags[e0 / kPerThread];
    if (fl) load_run<T>(tails, e0, ee, vals);
  }
  int64_t total;
  const int64_t x = block_excl_scan_i64<kBlock>((int64_t)__popc(fl), s_scan, &total);
  if (fl) {
    // Rows are consecutive and each row's out-list follows the previous one,
    // so the chunk's kept entries, in entry order, fill one contiguous range:
    // it starts in the chunk's first row u after u's kept entries in earlier
    // chunks (q[c] - q[first chunk of u]).
    const int64_t u = chunk_span(crows, c, nv).lo;
    const int64_t ou = off[u];
    const int64_t base = ooff[u] + (ou < eb ? q[c] - q[ou / kChunk] : 0) + x;
#pragma unroll
    for (int j = 0; j < kPerThread; j++)
      if ((fl >> j) & 1u) otails[base + __popc(fl & ((1u << j) - 1u))] = (int32_t)vals[j];
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan, reduce-then-scan over 4096-entry tiles (256 threads x 16).
constexpr int kScanItems = 16;
constexpr int kScanTile = 256 * kScanItems;

template <typename TIn>
__global__ __launch_bounds__(256) void k_scan_reduce(const TIn* __restrict__ in, int64_t n,
                                                     int64_t* __restrict__ bsum) {
  __shared__ uint64_t s_tmp[256 / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + k * 256 + threadIdx.x;
    if (i < n) acc += (uint64_t)in[i];
  }
  acc = block_sum_u64<256>(acc, s_tmp);
  if (threadIdx.x == 0) bsum[blockIdx.x] = (int64_t)acc;
}

// The tile goes through LDS (padded one slot in 32) so that global loads and
// stores stay coalesced while each thread scans 16 consecutive items.
template <typename TIn>
__global__ __launch_bounds__(256) void k_scan_apply(const TIn* __restrict__ in, int64_t n,
                                                    const int64_t* __restrict__ bpre,
                                                    int64_t* __restrict__ out) {
  __shared__ int64_t s_v[kScanTile + kScanTile / 32];
  __shared__ int64_t s_tmp[256 / kWave];
  auto pad = [](int i) { return i + (i >> 5); };
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int i = k * 256 + threadIdx.x;
    s_v[pad(i)] = base + i < n ? (int64_t)in[base + i] : 0;
  }
  __syncthreads();
  int64_t v[kScanItems];
  int64_t loc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = s_v[pad((int)threadIdx.x * kScanItems + k)];
    loc += v[k];
  }
  int64_t total;
  int64_t x = block_excl_scan_i64<256>(loc, s_tmp, &total);
  const int64_t pre = bpre ? bpre[blockIdx.x] : 0;
  x += pre;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    s_v[pad((int)threadIdx.x * kScanItems + k)] = x;  // this thread's own slots
    x += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int i = k * 256 + threadIdx.x;
    if (base + i < n) out[base + i] = s_v[pad(i)];
  }
  if (threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) out[n] = pre + total;
}

template <typename TIn>
static int scan_impl(xt_ctx* ctx, const TIn* in, int64_t n, int64_t* out, int64_t* tmp,
                     hipStream_t st) {
  if (n <= 0) {
    hipLaunchKernelGGL(k_fill<int64_t>, dim3(1), dim3(64), 0, st, out, (int64_t)1, (int64_t)0);
    XT_HIP(ctx, hipGetLastError());
    return XT_OK;
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply<TIn>, dim3(1), dim3(256), 0, st, in, n, (const int64_t*)nullptr,
                       out);
    XT_HIP(ctx, hipGetLastError());
    return XT_OK;
  }
  int64_t* bsum = tmp;
  int64_t* bpre = tmp + nb;
  int64_t* rest = bpre + nb + 1;
  hipLaunchKernelGGL(k_scan_reduce<TIn>, dim3((unsigned)nb), dim3(256), 0, st, in, n, bsum);
  XT_HIP(ctx, hipGetLastError());
  XT_TRY(scan_impl<int64_t>(ctx, bsum, nb, bpre, rest, st));
  hipLaunchKernelGGL(k_scan_apply<TIn>, dim3((unsigned)nb), dim3(256), 0, st, in, n,
                     (const int64_t*)bpre, out);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

static size_t scan_tmp_entries(int64_t n) {
  size_t tot = 0;
  while (n > kScanTile) {
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    tot += 2 * nb + 1;
    n = nb;
  }
  return tot + 8;
}

int scan_u32_to_i64(xt_ctx* ctx, Scratch& s, const uint32_t* in, int64_t n, int64_t* out,
                    hipStream_t st) {
  XT_TRY(ensure(ctx, s.scan_tmp, scan_tmp_entries(n) * sizeof(int64_t)));
  return scan_impl<uint32_t>(ctx, in, n, out, s.scan_tmp.as<int64_t>(), st);
}

int scan_u32_to_i64_tmp(xt_ctx* ctx, const uint32_t* in, int64_t n, int64_t* out, int64_t* tmp,
                        hipStream_t st) {
  return scan_impl<uint32_t>(ctx, in, n, out, tmp, st);
}

size_t scan_tmp_bytes(int64_t n) { return scan_tmp_entries(n) * sizeof(int64_t); }

static unsigned grid_for(int64_t n, int threads, unsigned cap = 65536u) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  return (unsigned)(g > cap ? cap : g);
}

int fill_u32(xt_ctx* ctx, uint32_t* p, int64_t n, uint32_t v, hipStream_t st) {
  if (n <= 0) return XT_OK;
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

int fill_u64(xt_ctx* ctx, uint64_t* p, int64_t n, uint64_t v, hipStream_t st) {
  if (n <= 0) return XT_OK;
  hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

int fill_i64(xt_ctx* ctx, int64_t* p, int64_t n, int64_t v, hipStream_t st) {
  if (n <= 0) return XT_OK;
  hipLaunchKernelGGL(k_fill<int64_t>, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

// ---------------------------------------------------------------------------
template <typename T>
int orient(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne, const int64_t* off, const T* tails,
           int mode, hipStream_t st) {
  XT_TRY(ensure(ctx, s.ooff, (size_t)(nv + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.otails, (size_t)(ne + kOtailsPad) * sizeof(int32_t)));
  XT_TRY(ensure(ctx, s.dplus, (size_t)(nv > 0 ? nv : 1) * sizeof(uint32_t)));
  if (nv <= 0) return XT_OK;
  if (ne <= 0) return fill_i64(ctx, s.ooff.as<int64_t>(), nv + 1, 0, st);
  const int64_t nchunks = (ne + kChunk - 1) / kChunk;
  XT_TRY(ensure(ctx, s.tailcnt, (size_t)nchunks * sizeof(uint32_t)));
  XT_TRY(ensure(ctx, s.qscan, (size_t)(nchunks + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.oflags, (size_t)(nchunks * kBlock)));
  XT_TRY(ensure(ctx, s.crows, (size_t)(2 * nchunks) * sizeof(int64_t)));
  int64_t* crows = s.crows.as<int64_t>();
  hipLaunchKernelGGL(k_chunk_rows, dim3(grid_for(nv, 256)), dim3(256), 0, st, nv, ne, off, crows);
  XT_HIP(ctx, hipGetLastError());
  uint16_t* deg = nullptr;
  if (mode == kOrientDegree) {
    XT_TRY(ensure(ctx, s.deg, (size_t)nv * sizeof(uint32_t)));
    deg = s.deg.as<uint16_t>();
    hipLaunchKernelGGL(k_degree, dim3(grid_for(nv, 256)), dim3(256), 0, st, nv, off, deg);
    XT_HIP(ctx, hipGetLastError());
  }
  XT_TRY(fill_u32(ctx, s.dplus.as<uint32_t>(), nv, 0u, st));
  uint64_t* counters = s.counters.as<uint64_t>();
  uint8_t* oflags = s.oflags.as<uint8_t>();
  if (mode == kOrientDegree)
    hipLaunchKernelGGL((k_orient_count<T, kOrientDegree>), dim3(xcd_grid(nchunks)), dim3(kBlock), 0,
                       st, nv, ne, off, tails, (const uint16_t*)deg, (const int64_t*)crows,
                       s.dplus.as<uint32_t>(), s.tailcnt.as<uint32_t>(), counters, oflags);
  else
    hipLaunchKernelGGL((k_orient_count<T, kOrientId>), dim3(xcd_grid(nchunks)), dim3(kBlock), 0, st,
                       nv, ne, off, tails, (const uint16_t*)deg, (const int64_t*)crows,
                       s.dplus.as<uint32_t>(), s.tailcnt.as<uint32_t>(), counters, oflags);
  XT_HIP(ctx, hipGetLastError());
  XT_TRY(scan_u32_to_i64(ctx, s, s.dplus.as<uint32_t>(), nv, s.ooff.as<int64_t>(), st));
  XT_TRY(scan_u32_to_i64(ctx, s, s.tailcnt.as<uint32_t>(), nchunks, s.qscan.as<int64_t>(), st));
  hipLaunchKernelGGL((k_orient_fill<T>), dim3(xcd_grid(nchunks)), dim3(kBlock), 0, st, nv, ne, off,
                     tails, (const uint8_t*)oflags, (const int64_t*)s.ooff.as<int64_t>(),
                     (const int64_t*)s.qscan.as<int64_t>(), (const int64_t*)crows, s.otails.as<int32_t>());
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

template int orient<int32_t>(xt_ctx*, Scratch&, int64_t, int64_t, const int64_t*, const int32_t*, int,
                             hipStream_t);
template int orient<int64_t>(xt_ctx*, Scratch&, int64_t, int64_t, const int64_t*, const int64_t*, int,
                             hipStream_t);

// ---------------------------------------------------------------------------
// Validation: ids in range, rows strictly ascending, optional symmetry.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_validate(int64_t nv, int64_t ne,
                                                     const int64_t* __restrict__ off,
                                                     const T* __restrict__ tails, int symmetry,
                                                     uint64_t* __restrict__ counters) {
  __shared__ int64_t s_off[kOffCap + 1];
  __shared__ int64_t s_lohi[2];
  const int64_t eb = (int64_t)blockIdx.x * kChunk;
  const int64_t ee = min(ne, eb + (int64_t)kChunk);
  const ChunkRows r = chunk_rows(off, nv, eb, ee, s_off, s_lohi);
  uint32_t bad_id = 0, bad_row = 0, asym = 0;
  for (int j = 0; j < kPerThread; j++) {
    const int64_t e = eb + (int64_t)j * kBlock + threadIdx.x;
    if (e >= ee) break;
    const int64_t u = chunk_row_of(r, s_off, off, e);
    const int64_t v = (int64_t)tails[e];
    if (v < 0 || v >= nv) {
      bad_id++;
      continue;
    }
    const int64_t rend = chunk_off(r, s_off, off, u + 1);
    if (e + 1 < rend && !((int64_t)tails[e + 1] > v)) bad_row++;
    if (symmetry) {
      const int64_t b0 = off[v], b1 = off[v + 1];
      const int64_t p = lower_bound_dev<T>(tails, b0, b1, (T)u);
      if (p >= b1 || (int64_t)tails[p] != u) asym++;
    }
  }
  if (bad_id) atomicAdd((unsigned long long*)&counters[kBadIds], (unsigned long long)bad_id);
  if (bad_row) atomicAdd((unsigned long long*)&counters[kBadRows], (unsigned long long)bad_row);
  if (asym) atomicAdd((unsigned long long*)&counters[kAsym], (unsigned long long)asym);
}

// Offsets check followed by the entry checks; reads the verdict back (syncs).
template <typename T>
int validate(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* off, const T* tails,
             bool symmetry, hipStream_t st, uint64_t* counters) {
  if (!counters) counters = ctx->s.counters.as<uint64_t>();
  XT_TRY(fill_u64(ctx, counters + kBadIds, 4, 0, st));
  hipLaunchKernelGGL(k_check_offsets, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, st, nv, ne, off,
                     counters);
  XT_HIP(ctx, hipGetLastError());
  uint64_t h[4];
  XT_HIP(ctx, hipMemcpyAsync(h, counters + kBadIds, sizeof(h), hipMemcpyDeviceToHost, st));
  XT_HIP(ctx, hipStreamSynchronize(st));
  if (h[kBadOff - kBadIds]) return set_err(ctx, XT_ERR_INPUT, "offsets are not a valid CSR index");
  if (ne > 0) {
    const int64_t nchunks = (ne + kChunk - 1) / kChunk;
    hipLaunchKernelGGL((k_validate<T>), dim3((unsigned)nchunks), dim3(kBlock), 0, st, nv, ne, off,
                       tails, symmetry ? 1 : 0, counters);
    XT_HIP(ctx, hipGetLastError());
    XT_HIP(ctx, hipMemcpyAsync(h, counters + kBadIds, sizeof(h), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
  }
  if (h[0]) return set_err(ctx, XT_ERR_INPUT, "neighbour id out of range [0, nv)");
  if (h[1]) return set_err(ctx, XT_ERR_INPUT, "adjacency row not strictly ascending (unsorted or duplicate entries)");
  if (h[2]) return set_err(ctx, XT_ERR_INPUT, "adjacency is not symmetric");
  return XT_OK;
}

template int validate<int32_t>(xt_ctx*, int64_t, int64_t, const int64_t*, const int32_t*, bool,
                               hipStream_t, uint64_t*);
template int validate<int64_t>(xt_ctx*, int64_t, int64_t, const int64_t*, const int64_t*, bool,
                               hipStream_t, uint64_t*);

}  // namespace xt
