// gen.hip -- seeded graph generators on the device (ingest, outside the timed count).
//
// R-MAT (scale s, edge factor f, probabilities a/b/c/d): tuple t in
// [0, f*2^s) descends s quadrant levels; level l draws 32 bits from
// mix64(key + 32*t + l/2) (low half for even l, high half for odd l) and
// compares them with the fixed-point thresholds floor(a*2^32),
// floor((a+b)*2^32), floor((a+b+c)*2^32).  Integer-only, so the host twin in
// the oracle reproduces it bit for bit.  Optional relabelling by a seeded
// bijection of [0, 2^s).  Cleaning: self-loops dropped, both directions kept,
// duplicates merged (radix sort of (u << s | v) keys, then unique).
//
// Erdos-Renyi G(n,p): edge {i<j} iff (mix64(key + i*n + j) >> 32) < thr64.
// One wave64 per row; a ballot per 64 candidates gives ordered compaction.
#include <rocprim/device/device_radix_sort.hpp>

#include "xt_device.hpp"
#include "xt_internal.hpp"

namespace xt {

int fill_u32(xt_ctx* ctx, uint32_t* p, int64_t n, uint32_t v, hipStream_t st);

__device__ __forceinline__ uint32_t rmat_perm_dev(uint32_t x, int scale, uint64_t key) {
  const uint64_t mask = (1ULL << scale) - 1;
  const int h = (scale + 1) / 2;
  uint64_t y = x;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const uint64_t m = mix64_dev(key + 1000 + r) | 1ULL;
    const uint64_t k = mix64_dev(key + 2000 + r);
    y = (y * m + k) & mask;
    y ^= y >> h;
  }
  return (uint32_t)y;
}

__global__ __launch_bounds__(256) void k_rmat_keys(int scale, int64_t nt, uint64_t key, uint32_t t0,
                                                   uint32_t t1, uint32_t t2, int permute,
                                                   uint64_t sentinel, uint64_t* __restrict__ keys) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
    uint32_t u = 0, v = 0;
    uint64_t w = 0;
    for (int l = 0; l < scale; l++) {
      if ((l & 1) == 0) w = mix64_dev(key + (uint64_t)t * 32 + (uint64_t)(l >> 1));
      const uint32_t r = (l & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
      const uint32_t ub = r >= t1 ? 1u : 0u;                       // quadrants c, d
      const uint32_t vb = (r >= t0 && r < t1) || r >= t2 ? 1u : 0u;  // quadrants b, d
      u = (u << 1) | ub;
      v = (v << 1) | vb;
    }
    if (permute) {
      u = rmat_perm_dev(u, scale, key);
      v = rmat_perm_dev(v, scale, key);
    }
    if (u == v) {
      keys[2 * t] = sentinel;
      keys[2 * t + 1] = sentinel;
    } else {
      keys[2 * t] = ((uint64_t)u << scale) | v;
      keys[2 * t + 1] = ((uint64_t)v << scale) | u;
    }
  }
}

__global__ __launch_bounds__(256) void k_mark_unique(const uint64_t* __restrict__ keys, int64_t n,
                                                     uint64_t sentinel, uint32_t* __restrict__ flags) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t k = keys[i];
    flags[i] = (k != sentinel && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
  }
}

__global__ __launch_bounds__(256) void k_scatter_unique(const uint64_t* __restrict__ keys, int64_t n,
                                                        const uint32_t* __restrict__ flags,
                                                        const int64_t* __restrict__ pos, int scale,
                                                        int32_t* __restrict__ tails,
                                                        uint32_t* __restrict__ deg) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const uint64_t mask = (1ULL << scale) - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (flags[i]) {
      const uint64_t k = keys[i];
      tails[pos[i]] = (int32_t)(k & mask);
      atomicAdd(&deg[k >> scale], 1u);
    }
  }
}

static unsigned grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  return (unsigned)(g > 65536 ? 65536 : g);
}

// Sorted, merged (u << bits | v) keys -> CSR: radix sort over key_bits bits,
// first of each run kept (the sentinel dropped), ids scattered to their rank,
// row lengths counted and scanned.  ka/kb hold nk+1 keys each (caller-owned;
// kb doubles as the rank buffer).
static int build_csr_from_keys(xt_ctx* ctx, uint64_t* ka, uint64_t* kb, int64_t nk, int key_bits, int bits,
                               int64_t nv, uint64_t sentinel, xt_graph* g) {
  hipStream_t st = ctx->stream;
  uint32_t* flags = nullptr;
  uint32_t* deg = nullptr;
  void* tmp = nullptr;
  int rc = XT_OK;
  auto cleanup = [&]() {
    if (flags) (void)hipFree(flags);
    if (deg) (void)hipFree(deg);
    if (tmp) (void)hipFree(tmp);
  };
#define GEN_HIP(expr)                                   \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) {                             \
      rc = hip_err(ctx, _e, #expr);                     \
      cleanup();                                        \
      return rc;                                        \
    }                                                   \
  } while (0)
  rocprim::double_buffer<uint64_t> db(ka, kb);
  size_t tmp_bytes = 0;
  GEN_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, (size_t)nk, 0u, (unsigned)key_bits, st));
  GEN_HIP(hipMalloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1));
  GEN_HIP(rocprim::radix_sort_keys(tmp, tmp_bytes, db, (size_t)nk, 0u, (unsigned)key_bits, st));
  (void)hipFree(tmp);
  tmp = nullptr;
  uint64_t* sorted = db.current();
  int64_t* pos = reinterpret_cast<int64_t*>(db.alternate());  // nk+1 entries, free after the sort
  GEN_HIP(hipMalloc(&flags, (size_t)(nk > 0 ? nk : 1) * sizeof(uint32_t)));
  if (nk > 0) {
    hipLaunchKernelGGL(k_mark_unique, dim3(grid_of(nk)), dim3(256), 0, st, sorted, nk, sentinel, flags);
    GEN_HIP(hipGetLastError());
  }
  rc = scan_u32_to_i64(ctx, ctx->s, flags, nk, pos, st);
  if (rc != XT_OK) {
    cleanup();
    return rc;
  }
  int64_t ne = 0;
  GEN_HIP(hipMemcpyAsync(&ne, pos + nk, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GEN_HIP(hipStreamSynchronize(st));
  g->nv = nv;
  g->ne = ne;
  GEN_HIP(hipMalloc(&g->off, (size_t)(nv + 1) * sizeof(int64_t)));
  GEN_HIP(hipMalloc(&g->tails, (size_t)(ne > 0 ? ne : 1) * sizeof(int32_t)));
  GEN_HIP(hipMalloc(&deg, (size_t)nv * sizeof(uint32_t)));
  rc = fill_u32(ctx, deg, nv, 0u, st);
  if (rc == XT_OK && nk > 0) {
    hipLaunchKernelGGL(k_scatter_unique, dim3(grid_of(nk)), dim3(256), 0, st, (const uint64_t*)sorted,
                       nk, (const uint32_t*)flags, (const int64_t*)pos, bits, g->tails, deg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_err(ctx, e, "k_scatter_unique");
  }
  if (rc == XT_OK) rc = scan_u32_to_i64(ctx, ctx->s, deg, nv, g->off, st);
  if (rc == XT_OK) {
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_err(ctx, e, "hipStreamSynchronize");
  }
  cleanup();
#undef GEN_HIP
  return rc;
}

int gen_rmat(xt_ctx* ctx, int scale, int edgefactor, const uint32_t thr[3], uint64_t seed,
             int permute, xt_graph* g) {
  if (scale < 1 || scale > 31 || edgefactor < 1) return set_err(ctx, XT_ERR_ARG, "R-MAT scale must be in [1,31]");
  hipStream_t st = ctx->stream;
  const int64_t nv = 1LL << scale;
  const int64_t nt = (int64_t)edgefactor * nv;
  const int64_t nk = 2 * nt;
  const uint64_t sentinel = (1ULL << (2 * scale)) - 1;  // (2^s-1, 2^s-1): a self-loop, never a real edge
  const uint64_t key = mix64(seed);
  uint64_t *ka = nullptr, *kb = nullptr;
  hipError_t e = hipMalloc(&ka, (size_t)(nk + 1) * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&kb, (size_t)(nk + 1) * sizeof(uint64_t));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rmat_keys, dim3(grid_of(nt)), dim3(256), 0, st, scale, nt, key, thr[0], thr[1],
                       thr[2], permute, sentinel, ka);
    e = hipGetLastError();
  }
  const int rc = e == hipSuccess ? build_csr_from_keys(ctx, ka, kb, nk, 2 * scale, scale, nv, sentinel, g)
                                 : hip_err(ctx, e, "R-MAT keys");
  if (ka) (void)hipFree(ka);
  if (kb) (void)hipFree(kb);
  return rc;
}

// ---------------------------------------------------------------------------
// Edge list -> symmetric CSR on the device: each pair (u, v) becomes the keys
// (u << bits | v) and (v << bits | u); self-loops become the sentinel (the
// self-loop of the largest representable id); ids outside [0, nv) are counted.
__global__ __launch_bounds__(256) void k_edge_keys(int64_t m, const int64_t* __restrict__ src,
                                                   const int64_t* __restrict__ dst, int64_t nv, int bits,
                                                   uint64_t sentinel, uint64_t* __restrict__ keys,
                                                   unsigned long long* __restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t nbad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const int64_t u = src[i], v = dst[i];
    const bool ok = u >= 0 && u < nv && v >= 0 && v < nv;
    nbad += ok ? 0u : 1u;
    if (!ok || u == v) {
      keys[2 * i] = sentinel;
      keys[2 * i + 1] = sentinel;
    } else {
      keys[2 * i] = ((uint64_t)u << bits) | (uint64_t)v;
      keys[2 * i + 1] = ((uint64_t)v << bits) | (uint64_t)u;
    }
  }
  if (nbad) atomicAdd(bad, (unsigned long long)nbad);
}

int graph_from_edges(xt_ctx* ctx, int64_t nv, int64_t m, const int64_t* src, const int64_t* dst, xt_graph* g) {
  hipStream_t st = ctx->stream;
  int bits = 1;
  while ((1ll << bits) < nv) bits++;
  const uint64_t sentinel = (1ULL << (2 * bits)) - 1;
  const int64_t nk = 2 * m;
  int64_t *ds = nullptr, *dd = nullptr;
  uint64_t *ka = nullptr, *kb = nullptr;
  unsigned long long* bad = nullptr;
  int rc = XT_OK;
  hipError_t e = hipMalloc(&ka, (size_t)(nk + 1) * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&kb, (size_t)(nk + 1) * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&bad, sizeof(unsigned long long));
  if (e == hipSuccess && m > 0) e = hipMalloc(&ds, (size_t)m * sizeof(int64_t));
  if (e == hipSuccess && m > 0) e = hipMalloc(&dd, (size_t)m * sizeof(int64_t));
  if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(unsigned long long), st);
  if (e == hipSuccess && m > 0) e = hipMemcpyAsync(ds, src, (size_t)m * sizeof(int64_t), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && m > 0) e = hipMemcpyAsync(dd, dst, (size_t)m * sizeof(int64_t), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && m > 0) {
    hipLaunchKernelGGL(k_edge_keys, dim3(grid_of(m)), dim3(256), 0, st, m, (const int64_t*)ds,
                       (const int64_t*)dd, nv, bits, sentinel, ka, bad);
    e = hipGetLastError();
  }
  unsigned long long nbad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess)
    rc = hip_err(ctx, e, "edge list upload");
  else if (nbad)
    rc = set_err(ctx, XT_ERR_INPUT, "edge list id out of range [0, nv)");
  if (ds) (void)hipFree(ds);
  if (dd) (void)hipFree(dd);
  if (bad) (void)hipFree(bad);
  if (rc == XT_OK) rc = build_csr_from_keys(ctx, ka, kb, nk, 2 * bits, bits, nv, sentinel, g);
  if (ka) (void)hipFree(ka);
  if (kb) (void)hipFree(kb);
  return rc;
}

// ---------------------------------------------------------------------------
// A resident CSR made a simple undirected graph: every entry (u, v) becomes the
// keys (u << bits | v) and (v << bits | u), so rows come out sorted, duplicates
// merged and missing reverse entries added; self-loops become the sentinel and
// are dropped.  One wave per row (a hub row spreads over the wave's lanes).
__global__ __launch_bounds__(256) void k_csr_keys(int64_t nv, const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ tails, int bits,
                                                  uint64_t sentinel, uint64_t* __restrict__ keys) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; u < nv; u += nw) {
    const int64_t a = off[u], b = off[u + 1];
    for (int64_t e = a + lane; e < b; e += kWave) {
      const int64_t v = tails[e];
      if (v == u) {
        keys[2 * e] = sentinel;
        keys[2 * e + 1] = sentinel;
      } else {
        keys[2 * e] = ((uint64_t)u << bits) | (uint64_t)v;
        keys[2 * e + 1] = ((uint64_t)v << bits) | (uint64_t)u;
      }
    }
  }
}

int graph_clean(xt_ctx* ctx, xt_graph* g) {
  hipStream_t st = ctx->stream;
  const int64_t nv = g->nv, ne = g->ne;
  int bits = 1;
  while ((1ll << bits) < nv) bits++;
  const uint64_t sentinel = (1ULL << (2 * bits)) - 1;
  const int64_t nk = 2 * ne;
  uint64_t *ka = nullptr, *kb = nullptr;
  hipError_t e = hipMalloc(&ka, (size_t)(nk + 1) * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&kb, (size_t)(nk + 1) * sizeof(uint64_t));
  if (e == hipSuccess && nv > 0) {
    hipLaunchKernelGGL(k_csr_keys, dim3(grid_of(nv * kWave)), dim3(256), 0, st, nv, (const int64_t*)g->off,
                       (const int32_t*)g->tails, bits, sentinel, ka);
    e = hipGetLastError();
  }
  xt_graph out;
  int rc = e == hipSuccess ? build_csr_from_keys(ctx, ka, kb, nk, 2 * bits, bits, nv, sentinel, &out)
                           : hip_err(ctx, e, "graph clean keys");
  if (ka) (void)hipFree(ka);
  if (kb) (void)hipFree(kb);
  if (rc != XT_OK) {
    if (out.off) (void)hipFree(out.off);
    if (out.tails) (void)hipFree(out.tails);
    return rc;
  }
  (void)hipFree(g->off);
  (void)hipFree(g->tails);
  g->off = out.off;
  g->tails = out.tails;
  g->ne = out.ne;
  return XT_OK;
}

// ---------------------------------------------------------------------------
// G(n,p): pass 0 counts each row, pass 1 writes it (ordered by ballot rank).
template <int PASS>
__global__ __launch_bounds__(256) void k_er_rows(int64_t n, uint64_t key, uint64_t thr64,
                                                 uint32_t* __restrict__ deg,
                                                 const int64_t* __restrict__ off,
                                                 int32_t* __restrict__ tails) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kWave;
  if (i >= n) return;  // whole wave leaves together
  uint32_t cnt = 0;
  int64_t out = PASS == 1 ? off[i] : 0;
  for (int64_t j0 = 0; j0 < n; j0 += kWave) {
    const int64_t j = j0 + lane;
    bool e = false;
    if (j < n && j != i) {
      const uint64_t a = i < j ? i : j, b = i < j ? j : i;
      e = (mix64_dev(key + a * (uint64_t)n + b) >> 32) < thr64;
    }
    const uint64_t m = __ballot(e);
    if constexpr (PASS == 1) {
      if (e) tails[out + __popcll(m & ((1ULL << lane) - 1))] = (int32_t)j;
      out += __popcll(m);
    } else {
      cnt += (uint32_t)__popcll(m);
    }
  }
  if constexpr (PASS == 0) {
    if (lane == 0) deg[i] = cnt;
  }
}

// Two-call G(n,p) into caller-provided device arrays: with tails == nullptr
// the rows are counted and off[0..n] is written (ne returned); with tails the
// rows are written using that off.  The degree scratch is s.deg.
int gen_er_rows(xt_ctx* ctx, Scratch& s, int64_t n, uint64_t seed, uint64_t thr64, int64_t* off,
                int32_t* tails, int64_t cap, int64_t* ne_out, hipStream_t st) {
  if (n < 1 || n > (1LL << 31) - 1) return set_err(ctx, XT_ERR_ARG, "ER n out of range");
  const uint64_t key = mix64(seed);
  const unsigned blocks = (unsigned)((n + 3) / 4);
  if (!tails) {
    XT_TRY(ensure(ctx, s.deg, (size_t)n * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_er_rows<0>, dim3(blocks), dim3(256), 0, st, n, key, thr64, s.deg.as<uint32_t>(),
                       (const int64_t*)nullptr, (int32_t*)nullptr);
    XT_HIP(ctx, hipGetLastError());
    XT_TRY(scan_u32_to_i64(ctx, s, s.deg.as<uint32_t>(), n, off, st));
    int64_t ne = 0;
    XT_HIP(ctx, hipMemcpyAsync(&ne, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
    if (ne_out) *ne_out = ne;
    return XT_OK;
  }
  if (ne_out && *ne_out > cap) return set_err(ctx, XT_ERR_CAPACITY, "ER graph exceeds the tails capacity");
  hipLaunchKernelGGL(k_er_rows<1>, dim3(blocks), dim3(256), 0, st, n, key, thr64, (uint32_t*)nullptr,
                     (const int64_t*)off, tails);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

int gen_er(xt_ctx* ctx, int64_t n, uint64_t seed, uint64_t thr64, xt_graph* g) {
  if (n < 1 || n > (1LL << 31) - 1) return set_err(ctx, XT_ERR_ARG, "ER n out of range");
  hipStream_t st = ctx->stream;
  g->nv = n;
  XT_HIP(ctx, hipMalloc(&g->off, (size_t)(n + 1) * sizeof(int64_t)));
  int64_t ne = 0;
  XT_TRY(gen_er_rows(ctx, ctx->s, n, seed, thr64, g->off, nullptr, 0, &ne, st));
  g->ne = ne;
  XT_HIP(ctx, hipMalloc(&g->tails, (size_t)(ne > 0 ? ne : 1) * sizeof(int32_t)));
  XT_TRY(gen_er_rows(ctx, ctx->s, n, seed, thr64, g->off, g->tails, ne, &ne, st));
  XT_HIP(ctx, hipStreamSynchronize(st));
  return XT_OK;
}

// ---------------------------------------------------------------------------
// One rank's slice of a 1-D vertex split (graph.hpp:76-80 splits vertices into
// contiguous blocks): the rows of the owned vertices [lo, hi) and of every
// neighbour of them (ghosts) kept whole, all other rows emptied; ids stay
// global.  Every triangle whose minimum vertex is owned has all three edges
// in it: two in the owned row, the third in a ghost row.
__global__ __launch_bounds__(256) void k_slice_mark(int64_t lo, int64_t hi, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ tails,
                                                    uint8_t* __restrict__ mark) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t e0 = off[lo], e1 = off[hi];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (hi - lo) + (e1 - e0); i += stride) {
    if (i < hi - lo)
      mark[lo + i] = 1;  // owned
    else
      mark[tails[e0 + (i - (hi - lo))]] = 1;  // neighbour of an owned vertex
  }
}

__global__ __launch_bounds__(256) void k_slice_degrees(int64_t nv, const int64_t* __restrict__ off,
                                                       const uint8_t* __restrict__ mark,
                                                       uint32_t* __restrict__ deg) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nv; u += stride)
    deg[u] = mark[u] ? (uint32_t)(off[u + 1] - off[u]) : 0u;
}

// One wave per row; the lanes copy the row's ids.
__global__ __launch_bounds__(256) void k_slice_copy(int64_t nv, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ tails,
                                                    const uint8_t* __restrict__ mark,
                                                    const int64_t* __restrict__ noff,
                                                    int32_t* __restrict__ ntails) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; u < nv; u += nw) {
    if (!mark[u]) continue;
    const int64_t a = off[u], d = off[u + 1] - a, b = noff[u];
    for (int64_t i = lane; i < d; i += kWave) ntails[b + i] = tails[a + i];
  }
}

int graph_slice(xt_ctx* ctx, const xt_graph* g, int64_t lo, int64_t hi, xt_graph* out) {
  const int64_t nv = g->nv;
  hipStream_t st = ctx->stream;
  uint8_t* mark = nullptr;
  uint32_t* deg = nullptr;
  int64_t* tmp = nullptr;
  int rc = XT_OK;
  auto fail = [&](hipError_t e, const char* what) {
    rc = hip_err(ctx, e, what);
    return rc;
  };
  hipError_t e = hipMalloc(&mark, (size_t)(nv > 0 ? nv : 1));
  if (e == hipSuccess) e = hipMalloc(&deg, (size_t)(nv > 0 ? nv : 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&tmp, scan_tmp_bytes(nv > 0 ? nv : 1));
  if (e == hipSuccess) e = hipMalloc(&out->off, (size_t)(nv + 1) * sizeof(int64_t));
  if (e != hipSuccess) fail(e, "hipMalloc(slice)");
  if (rc == XT_OK && nv > 0) {
    e = hipMemsetAsync(mark, 0, (size_t)nv, st);
    if (e != hipSuccess) fail(e, "hipMemsetAsync(slice)");
  }
  if (rc == XT_OK && nv > 0) {
    hipLaunchKernelGGL(k_slice_mark, dim3(grid_of(nv)), dim3(256), 0, st, lo, hi, (const int64_t*)g->off,
                       (const int32_t*)g->tails, mark);
    hipLaunchKernelGGL(k_slice_degrees, dim3(grid_of(nv)), dim3(256), 0, st, nv, (const int64_t*)g->off,
                       (const uint8_t*)mark, deg);
    e = hipGetLastError();
    if (e != hipSuccess) fail(e, "slice kernels");
  }
  if (rc == XT_OK) rc = scan_u32_to_i64_tmp(ctx, deg, nv, out->off, tmp, st);
  int64_t ne = 0;
  if (rc == XT_OK) {
    e = hipMemcpyAsync(&ne, out->off + nv, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) fail(e, "slice size");
  }
  if (rc == XT_OK) {
    e = hipMalloc(&out->tails, (size_t)(ne > 0 ? ne : 1) * sizeof(int32_t));
    if (e != hipSuccess) fail(e, "hipMalloc(slice tails)");
  }
  if (rc == XT_OK && nv > 0) {
    hipLaunchKernelGGL(k_slice_copy, dim3(grid_of(nv * kWave)), dim3(256), 0, st, nv, (const int64_t*)g->off,
                       (const int32_t*)g->tails, (const uint8_t*)mark, (const int64_t*)out->off, out->tails);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) fail(e, "slice copy");
  }
  out->nv = nv;
  out->ne = ne;
  if (mark) (void)hipFree(mark);
  if (deg) (void)hipFree(deg);
  if (tmp) (void)hipFree(tmp);
  return rc;
}

}  // namespace xt
