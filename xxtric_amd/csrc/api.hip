// api.hip -- extern "C" entry points of libxtric (context, single-graph counts,
// reference lower-FFI shapes, resident graphs).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "xt_device.hpp"
#include "xt_internal.hpp"

namespace xt {

int launch_masked(xt_ctx* ctx, int64_t nedges, const int64_t* mask, const int64_t* ind,
                  const int64_t* L, int64_t* out, hipStream_t st);
int launch_inner(xt_ctx* ctx, int64_t start, int64_t end, int64_t jstart, int64_t jend,
                 const int64_t* L, int64_t* out, hipStream_t st);
int launch_add_total(xt_ctx* ctx, int64_t* out, hipStream_t st, const uint64_t* counters = nullptr);

__global__ __launch_bounds__(256) void k_copy16(int4* __restrict__ dst, const int4* __restrict__ src,
                                                int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

static int launch_copy16(xt_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st) {
  const int64_t n = (int64_t)(bytes / 16);
  if (n == 0) return XT_OK;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)(ctx->num_cus * 16)), dim3(256), 0, st,
                     static_cast<int4*>(dst), static_cast<const int4*>(src), n);
  XT_HIP(ctx, hipGetLastError());
  return XT_OK;
}

// ---- memory ceilings: the read paths the counting kernels are bound by ------
// Streaming read: every 16-byte element once, coalesced (the group kernel's
// streamed lists).  Random gather: one elem-byte load per index from a
// pseudo-random position (the orientation and light passes' neighbour keys
// and records).  Each thread folds what it loads into one word, written once,
// so no load is dead.
__global__ __launch_bounds__(256) void k_stream_read(const int4* __restrict__ a, int64_t n,
                                                     uint32_t* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t acc = 0;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 7 * stride < n; i += 8 * stride) {  // eight coalesced 16-byte loads in flight
    int4 v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = a[i + j * stride];
#pragma unroll
    for (int j = 0; j < 8; j++) acc ^= (uint32_t)(v[j].x ^ v[j].y ^ v[j].z ^ v[j].w);
  }
  for (; i < n; i += stride) {
    const int4 v = a[i];
    acc ^= (uint32_t)(v.x ^ v.y ^ v.z ^ v.w);
  }
  sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <typename E>
__global__ __launch_bounds__(256) void k_gather(const E* __restrict__ a, int64_t n, int64_t ngather,
                                                uint32_t* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t acc = 0;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 7 * stride < ngather; i += 8 * stride) {  // eight independent gathers in flight
    E v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = a[mix64_dev((uint64_t)(i + j * stride)) & (uint64_t)(n - 1)];
#pragma unroll
    for (int j = 0; j < 8; j++) acc ^= (uint32_t)v[j];
  }
  for (; i < ngather; i += stride) acc ^= (uint32_t)a[mix64_dev((uint64_t)i) & (uint64_t)(n - 1)];
  sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

int set_err(xt_ctx* ctx, int code, const std::string& msg) {
  if (ctx) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);  // the server's submitters set errors too
    ctx->err = msg;
  }
  return code;
}

int hip_err(xt_ctx* ctx, hipError_t e, const char* what) {
  const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? XT_ERR_NOMEM : XT_ERR_HIP;
  return set_err(ctx, code, std::string(what) + ": " + hipGetErrorString(e));
}

int ensure(xt_ctx* ctx, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return XT_OK;
  if (b.fixed)
    return set_err(ctx, XT_ERR_CAPACITY, std::string("fixed scratch slice of ") + std::to_string(b.bytes) +
                                             " bytes asked for " + std::to_string(bytes));
  if (b.p) {
    // Work queued on any stream (a caller's stream included) may still read it.
    XT_HIP(ctx, hipDeviceSynchronize());
    XT_HIP(ctx, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  const size_t want = (bytes + ((size_t)1 << 21) - 1) & ~(((size_t)1 << 21) - 1);
  if (ctx && ctx->fail_allocs > 0) {
    ctx->fail_allocs--;
    return set_err(ctx, XT_ERR_NOMEM, std::string("device allocation of ") + std::to_string(want) +
                                          " bytes failed (injected)");
  }
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return set_err(ctx, XT_ERR_NOMEM, std::string("device allocation of ") + std::to_string(want) +
                                          " bytes failed: " + hipGetErrorString(e));
  }
  b.bytes = want;
  return XT_OK;
}

// ---- scratch of the allocation-free entry point ----------------------------
// One statically sized device array, resident on every GPU the library runs
// on until the process ends.  Its host-side name is not a device address:
// host code uses only the per-device address from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocDevice {
  std::mutex init_mu;   // guards ptr / ready
  bool ready = false;
  unsigned char* ptr = nullptr;  // this device's copy of `scratch`
  std::mutex call_mu;   // one allocation-free count at a time on this device
};
static NoAllocDevice g_noalloc[kMaxDevices];

// Looks up this device's address of `scratch` once; the device must be current.
static int noalloc_init(xt_ctx* ctx, int device) {
  if (device < 0 || device >= kMaxDevices) return set_err(ctx, XT_ERR_ARG, "device ordinal beyond 63");
  NoAllocDevice& d = g_noalloc[device];
  std::lock_guard<std::mutex> lk(d.init_mu);
  if (d.ready) return XT_OK;
  void* p = nullptr;
  XT_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return set_err(ctx, XT_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
  d.ptr = static_cast<unsigned char*>(p);
  d.ready = true;
  return XT_OK;
}

static int check_host_offsets(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* off) {
  if (nv < 0 || ne < 0 || !off) return set_err(ctx, XT_ERR_ARG, "negative size or null offsets");
  if (nv >= (1LL << 31) - 1) return set_err(ctx, XT_ERR_ARG, "nv must be below 2^31-1 (int32 device ids)");
  if (off[0] != 0 || off[nv] != ne) return set_err(ctx, XT_ERR_INPUT, "offsets must start at 0 and end at ne");
  for (int64_t i = 0; i < nv; i++)
    if (off[i] > off[i + 1]) return set_err(ctx, XT_ERR_INPUT, "offsets are not monotone");
  return XT_OK;
}

template <typename T>
int count_device_csr(xt_ctx* ctx, Scratch& s, int64_t nv, int64_t ne, const int64_t* off,
                     const T* tails, uint32_t flags, hipStream_t st, bool timed, const OwnedRange* range) {
  const int mode = (flags & XT_ORIENT_ID) ? kOrientId : kOrientDegree;
  XT_TRY(ensure(ctx, s.counters, kNumSlots * sizeof(uint64_t)));
  XT_TRY(fill_u64(ctx, s.counters.as<uint64_t>(), kNumSlots, 0, st));
  uint64_t* pv = nullptr;
  if (flags & XT_PER_VERTEX) {
    XT_TRY(ensure(ctx, s.pv, (size_t)(nv > 0 ? nv : 1) * sizeof(uint64_t)));
    pv = s.pv.as<uint64_t>();
  }
  // Grow every scratch buffer before the first event so no allocation (and
  // its implicit synchronisation) falls inside the timed window.
  XT_TRY(ensure(ctx, s.ooff, (size_t)(nv + 1) * sizeof(int64_t)));
  XT_TRY(ensure(ctx, s.otails, (size_t)(ne + kOtailsPad) * sizeof(int32_t)));
  if (timed) XT_HIP(ctx, hipEventRecord(ctx->ev[kEvStart], st));
  XT_TRY(orient<T>(ctx, s, nv, ne, off, tails, mode, st));
  if (timed) XT_HIP(ctx, hipEventRecord(ctx->ev[kEvOriented], st));
  if (pv) XT_TRY(fill_u64(ctx, pv, nv, 0, st));
  XT_TRY(count_oriented<T>(ctx, s, nv, ne, off, tails, pv, nullptr, nullptr, st,
                        timed ? &ctx->ev[kEvLight] : nullptr, range));
  if (timed) XT_HIP(ctx, hipEventRecord(ctx->ev[kEvCounted], st));
  return XT_OK;
}

template int count_device_csr<int32_t>(xt_ctx*, Scratch&, int64_t, int64_t, const int64_t*, const int32_t*,
                                       uint32_t, hipStream_t, bool, const OwnedRange*);
template int count_device_csr<int64_t>(xt_ctx*, Scratch&, int64_t, int64_t, const int64_t*, const int64_t*,
                                       uint32_t, hipStream_t, bool, const OwnedRange*);

// Reads the count, the input-error slots and the timing after a count.
static int finish_count(xt_ctx* ctx, int64_t nv, uint32_t flags, hipStream_t st, bool timed,
                        uint64_t* total, uint64_t* per_vertex) {
  uint64_t h[kNumSlots];
  XT_HIP(ctx, hipMemcpyAsync(h, ctx->s.counters.p, sizeof(h), hipMemcpyDeviceToHost, st));
  if (per_vertex && (flags & XT_PER_VERTEX) && nv > 0)
    XT_HIP(ctx, hipMemcpyAsync(per_vertex, ctx->s.pv.p, (size_t)nv * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, st));
  XT_HIP(ctx, hipStreamSynchronize(st));
  if (h[kBadIds]) return set_err(ctx, XT_ERR_INPUT, "neighbour id out of range [0, nv)");
  if (total) *total = h[kTotal];
  int64_t nep = 0, nheavy = 0, ntiles = 0, nwave = 0;
  const bool worked = nv > 0 && ctx->s.itcnt.p != nullptr && ctx->s.ooff.p != nullptr;
  if (nv > 0)
    XT_HIP(ctx, hipMemcpy(&nep, ctx->s.ooff.as<int64_t>() + nv, sizeof(int64_t), hipMemcpyDeviceToHost));
  const bool split = worked && nep > 0;  // the light/items/group phases ran
  if (split) {
    nheavy = (int64_t)h[kHeavyItems];
    XT_HIP(ctx, hipMemcpy(&ntiles, ctx->s.tileoff.as<int64_t>() + nv, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (ctx->last_wave_owner > 0) {
      int64_t wend = 0;
      XT_HIP(ctx, hipMemcpy(&wend, ctx->s.tileoff.as<int64_t>() + 2 * nv, sizeof(int64_t), hipMemcpyDeviceToHost));
      nwave = wend - ntiles;
    }
  }
  ctx->timing = xt_timing{};
  ctx->timing.ne_plus = nep;
  ctx->timing.n_heavy = nheavy;
  ctx->timing.n_tiles = ntiles;
  ctx->timing.n_wave_tiles = nwave;
  if (timed) {
    float a = 0, b = 0, c = 0;
    XT_HIP(ctx, hipEventElapsedTime(&a, ctx->ev[kEvStart], ctx->ev[kEvOriented]));
    XT_HIP(ctx, hipEventElapsedTime(&b, ctx->ev[kEvOriented], ctx->ev[kEvCounted]));
    XT_HIP(ctx, hipEventElapsedTime(&c, ctx->ev[kEvStart], ctx->ev[kEvCounted]));
    ctx->timing.orient_ms = a;
    ctx->timing.count_ms = b;
    ctx->timing.total_ms = c;
    if (split) {
      float l = 0, i = 0, g = 0;
      XT_HIP(ctx, hipEventElapsedTime(&l, ctx->ev[kEvOriented], ctx->ev[kEvLight]));
      XT_HIP(ctx, hipEventElapsedTime(&i, ctx->ev[kEvLight], ctx->ev[kEvItems]));
      XT_HIP(ctx, hipEventElapsedTime(&g, ctx->ev[kEvItems], ctx->ev[kEvCounted]));
      ctx->timing.light_ms = l;
      ctx->timing.items_ms = i;
      ctx->timing.group_ms = g;
    }
  }
  return XT_OK;
}

// Host int64 ids -> int32 for the device layout.
static void narrow_ids(const int64_t* in, int32_t* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = (int32_t)in[i];
}

static int upload_csr(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* off, const void* tails,
                      int tails_bytes, int64_t* d_off, int32_t* d_tails, hipStream_t st) {
  XT_HIP(ctx, hipMemcpyAsync(d_off, off, (size_t)(nv + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (ne == 0) return XT_OK;
  if (tails_bytes == 4) {
    XT_HIP(ctx, hipMemcpyAsync(d_tails, tails, (size_t)ne * sizeof(int32_t), hipMemcpyHostToDevice, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
    return XT_OK;
  }
  const int64_t* t64 = static_cast<const int64_t*>(tails);
  for (int64_t e = 0; e < ne; e++)
    if (t64[e] < 0 || t64[e] >= nv) return set_err(ctx, XT_ERR_INPUT, "neighbour id out of range [0, nv)");
  std::vector<int32_t> tmp((size_t)ne);
  narrow_ids(t64, tmp.data(), ne);
  XT_HIP(ctx, hipMemcpyAsync(d_tails, tmp.data(), (size_t)ne * sizeof(int32_t), hipMemcpyHostToDevice, st));
  XT_HIP(ctx, hipStreamSynchronize(st));
  return XT_OK;
}

}  // namespace xt

using namespace xt;

extern "C" {

const char* xt_version(void) { return "xtric 0.1 (gfx950)"; }

int xt_ctx_create(int device, xt_ctx** out) {
  return xt_guarded([&]() -> int {
    if (!out) return XT_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return XT_ERR_HIP;
    if (device < 0 || device >= ndev) return XT_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return XT_ERR_HIP;
    xt_ctx* ctx = new xt_ctx();
    ctx->device = device;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
      ctx->num_cus = ncu;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return XT_ERR_HIP;
    }
    for (int i = 0; i < kNumEv; i++)
      if (hipEventCreate(&ctx->ev[i]) != hipSuccess) {
        delete ctx;
        return XT_ERR_HIP;
      }
    if (ensure(ctx, ctx->s.counters, kNumSlots * sizeof(uint64_t)) != XT_OK) {
      release(ctx->s);
      delete ctx;
      return XT_ERR_NOMEM;
    }
    if (noalloc_init(ctx, device) != XT_OK) {
      release(ctx->s);
      delete ctx;
      return XT_ERR_HIP;
    }
    *out = ctx;
    return XT_OK;
  });
}

int xt_ctx_destroy(xt_ctx* ctx) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    release(ctx->s);
    for (int i = 0; i < kNumEv; i++)
      if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return XT_OK;
  });
}

const char* xt_last_error(const xt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int xt_last_timing(const xt_ctx* ctx, xt_timing* out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out) return XT_ERR_ARG;
    *out = ctx->timing;
    return XT_OK;
  });
}

int xt_ctx_tune(xt_ctx* ctx, int light_cut, int group_blocks, int group_round) {
  return xt_guarded([&]() -> int {
    if (!ctx || group_blocks > 4 || light_cut < -1 || group_blocks < 0) return XT_ERR_ARG;
    if (group_round != 0 && group_round != 1 && group_round != 2 && group_round != 4) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (light_cut > 0) ctx->small_cut = light_cut;
    if (light_cut == -1) ctx->small_cut = 0;  // back to the defaults (one graph / batches)
    if (group_blocks > 0) ctx->group_blocks = group_blocks;
    if (group_round > 0) ctx->group_round = group_round;
    return XT_OK;
  });
}

int xt_ctx_tune_tiles(xt_ctx* ctx, int items_per_tile, int tile_run) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    if (items_per_tile != 0 && items_per_tile != -1 && items_per_tile != 64 && items_per_tile != 128 &&
        items_per_tile != 256)
      return set_err(ctx, XT_ERR_ARG, "items_per_tile must be 64, 128 or 256 (0 keeps, -1 restores 256)");
    if (tile_run < -1 || tile_run > 64) return set_err(ctx, XT_ERR_ARG, "tile_run must be in 1..64 (0 keeps, -1 restores 16)");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (items_per_tile > 0) ctx->items_per_tile = items_per_tile;
    if (items_per_tile == -1) ctx->items_per_tile = 256;
    if (tile_run > 0) ctx->tile_run = tile_run;
    if (tile_run == -1) ctx->tile_run = 16;
    return XT_OK;
  });
}

int xt_ctx_tune_waves(xt_ctx* ctx, int wave_owner) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    if (wave_owner < -2 || wave_owner > 512)
      return set_err(ctx, XT_ERR_ARG, "wave_owner must be in 1..512 (0 keeps, -1 restores the defaults, -2 turns wave tiles off)");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (wave_owner > 0) ctx->wave_owner = wave_owner;
    if (wave_owner == -1) ctx->wave_owner = -1;
    if (wave_owner == -2) ctx->wave_owner = 0;
    return XT_OK;
  });
}

int xt_ctx_tune_hash(xt_ctx* ctx, int group_hash) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    if (group_hash < -1 || group_hash > 2)
      return set_err(ctx, XT_ERR_ARG, "group_hash must be 1 (multiplicative) or 2 (low bits); 0 keeps, -1 restores the default");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (group_hash > 0) ctx->group_hash = group_hash;
    if (group_hash == -1) ctx->group_hash = kHashMix;
    return XT_OK;
  });
}

int xt_ctx_fail_allocs(xt_ctx* ctx, int n) {
  return xt_guarded([&]() -> int {
    if (!ctx || n < 0) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->fail_allocs = n;
    return XT_OK;
  });
}

int xt_ctx_force_wide(xt_ctx* ctx, int on) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->force_wide = on != 0;
    return XT_OK;
  });
}

int xt_dev_alloc(xt_ctx* ctx, size_t bytes, void** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    XT_HIP(ctx, hipMalloc(out, bytes > 0 ? bytes : 1));
    return XT_OK;
  });
}

int xt_dev_free(xt_ctx* ctx, void* p) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (p) {
      XT_HIP(ctx, hipDeviceSynchronize());
      XT_HIP(ctx, hipFree(p));
    }
    return XT_OK;
  });
}

int xt_memcpy(xt_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
  return xt_guarded([&]() -> int {
    if (!ctx || (bytes > 0 && (!dst || !src))) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (kind == XT_D2D_KERNEL) {
      if (bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0)
        return set_err(ctx, XT_ERR_ARG, "kernel copy needs 16-byte aligned sizes and pointers");
      XT_TRY(launch_copy16(ctx, dst, src, bytes, ctx->stream));
      XT_HIP(ctx, hipStreamSynchronize(ctx->stream));
      return XT_OK;
    }
    hipMemcpyKind k = kind == XT_H2D ? hipMemcpyHostToDevice
                      : kind == XT_D2H ? hipMemcpyDeviceToHost
                      : kind == XT_D2D ? hipMemcpyDeviceToDevice
                                       : hipMemcpyDefault;
    if (kind < XT_H2D || kind > XT_D2D_KERNEL) return set_err(ctx, XT_ERR_ARG, "bad copy kind");
    if (bytes == 0) return XT_OK;
    XT_HIP(ctx, hipDeviceSynchronize());
    XT_HIP(ctx, hipMemcpy(dst, src, bytes, k));
    return XT_OK;
  });
}

int xt_mem_ceiling(xt_ctx* ctx, int kind, int64_t n, int elem_bytes, int64_t ngather, int reps,
                   double* ms_out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !ms_out || n <= 0 || reps < 1) return XT_ERR_ARG;
    if (kind == 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
      return set_err(ctx, XT_ERR_ARG, "gather elements are 2, 4 or 8 bytes");
    if (kind == 1 && (n & (n - 1)) != 0) return set_err(ctx, XT_ERR_ARG, "gather arrays hold a power of two elements");
    if (kind != 0 && kind != 1) return set_err(ctx, XT_ERR_ARG, "kind 0 (streaming read) or 1 (gather)");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = kind == 0 ? (size_t)n * 16 : (size_t)n * (size_t)elem_bytes;
    const unsigned grid = (unsigned)(ctx->num_cus * 16);
    void* a = nullptr;
    uint32_t* sink = nullptr;
    XT_HIP(ctx, hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&sink, (size_t)grid * 256 * sizeof(uint32_t));
    int rc = XT_OK;
    if (e == hipSuccess) e = hipMemsetAsync(a, 0x5a, bytes, ctx->stream);
    float best = 0.f;
    for (int r = 0; e == hipSuccess && r <= reps; r++) {  // run 0 warms up
      XT_HIP(ctx, hipEventRecord(ctx->ev[kEvStart], ctx->stream));
      if (kind == 0)
        hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, ctx->stream, (const int4*)a, n, sink);
      else if (elem_bytes == 2)
        hipLaunchKernelGGL(k_gather<uint16_t>, dim3(grid), dim3(256), 0, ctx->stream, (const uint16_t*)a, n,
                           ngather, sink);
      else if (elem_bytes == 4)
        hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream, (const uint32_t*)a, n,
                           ngather, sink);
      else
        hipLaunchKernelGGL(k_gather<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream, (const uint64_t*)a, n,
                           ngather, sink);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(ctx->ev[kEvCounted], ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[kEvCounted]);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev[kEvStart], ctx->ev[kEvCounted]);
      if (r > 0 && (best == 0.f || ms < best)) best = ms;
    }
    if (e != hipSuccess) rc = hip_err(ctx, e, "memory ceiling kernels");
    (void)hipFree(a);
    if (sink) (void)hipFree(sink);
    *ms_out = best;
    return rc;
  });
}

int xt_dev_sync(xt_ctx* ctx) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    XT_HIP(ctx, hipSetDevice(ctx->device));
    XT_HIP(ctx, hipDeviceSynchronize());
    return XT_OK;
  });
}

int xt_count_csr(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* offsets, const int64_t* tails,
                 uint32_t flags, uint64_t* total, uint64_t* per_vertex) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (!total || (ne > 0 && !tails)) return set_err(ctx, XT_ERR_ARG, "null output or tails");
    if (per_vertex && !(flags & XT_PER_VERTEX)) return set_err(ctx, XT_ERR_ARG, "per_vertex needs XT_PER_VERTEX");
    XT_TRY(check_host_offsets(ctx, nv, ne, offsets));
    hipStream_t st = ctx->stream;
    XT_TRY(ensure(ctx, ctx->s.in_off, (size_t)(nv + 1) * sizeof(int64_t)));
    XT_TRY(ensure(ctx, ctx->s.in_tails, (size_t)(ne > 0 ? ne : 1) * sizeof(int32_t)));
    int64_t* d_off = ctx->s.in_off.as<int64_t>();
    int32_t* d_tails = ctx->s.in_tails.as<int32_t>();
    XT_TRY(upload_csr(ctx, nv, ne, offsets, tails, 8, d_off, d_tails, st));
    if (flags & XT_VALIDATE) XT_TRY(validate<int32_t>(ctx, nv, ne, d_off, d_tails, (flags & XT_CHECK_SYMMETRY) != 0, st));
    XT_TRY(count_device_csr<int32_t>(ctx, ctx->s, nv, ne, d_off, d_tails, flags, st, true));
    return finish_count(ctx, nv, flags, st, true, total, per_vertex);
  });
}

int xt_single_gpu_count(xt_ctx* ctx, int64_t nv, const int64_t* d_indices, const int64_t* d_edge_list,
                        int64_t* d_local_count, void* stream) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (nv < 0 || !d_indices || !d_local_count) return set_err(ctx, XT_ERR_ARG, "bad arguments");
    // Device ids are int32 past this point: reject before d_indices[nv] is read.
    if (nv >= (1LL << 31) - 1) return set_err(ctx, XT_ERR_ARG, "nv must be below 2^31-1 (int32 device ids)");
    if (nv == 0) return XT_OK;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    int64_t ne = 0;
    XT_HIP(ctx, hipMemcpyAsync(&ne, d_indices + nv, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
    if (ne < 0) return set_err(ctx, XT_ERR_INPUT, "negative edge count in d_indices[nv]");
    if (ne > 0 && !d_edge_list) return set_err(ctx, XT_ERR_ARG, "null d_edge_list");
    // Offsets must be a valid index (and rows sorted) before the chunked
    // kernels search them; the reference launcher assumes the same.
    XT_TRY(validate<int64_t>(ctx, nv, ne, d_indices, d_edge_list, false, st));
    XT_TRY(count_device_csr<int64_t>(ctx, ctx->s, nv, ne, d_indices, d_edge_list, XT_ORIENT_DEGREE, st, false));
    uint64_t bad = 0;
    XT_HIP(ctx, hipMemcpyAsync(&bad, ctx->s.counters.as<uint64_t>() + kBadIds, sizeof(bad),
                               hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
    if (bad) return set_err(ctx, XT_ERR_INPUT, "neighbour id out of range [0, nv)");
    XT_TRY(launch_add_total(ctx, d_local_count, st));
    // The add reads the context's counters, which the next call on this
    // context (on any stream) clears: it must be done before this returns.
    XT_HIP(ctx, hipStreamSynchronize(st));
    return XT_OK;
  });
}

int xt_single_gpu_count_noalloc(xt_ctx* ctx, int64_t nv, const int64_t* d_indices,
                                const int64_t* d_edge_list, int64_t* d_local_count, void* stream) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (nv < 0 || !d_indices || !d_local_count) return set_err(ctx, XT_ERR_ARG, "bad arguments");
    if (nv >= (1LL << 31) - 1) return set_err(ctx, XT_ERR_ARG, "nv must be below 2^31-1 (int32 device ids)");
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_noalloc[ctx->device].ready)
      return set_err(ctx, XT_ERR_STATE, "no scratch address for this device");
    NoAllocDevice& dev = g_noalloc[ctx->device];
    if (nv == 0) return XT_OK;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    int64_t ne = 0;
    XT_HIP(ctx, hipMemcpyAsync(&ne, d_indices + nv, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
    if (ne < 0) return set_err(ctx, XT_ERR_INPUT, "negative edge count in d_indices[nv]");
    if (ne > 0 && !d_edge_list) return set_err(ctx, XT_ERR_ARG, "null d_edge_list");
    const size_t need = carve_bytes(nv, ne);
    if (need > kNoAllocScratchBytes)
      return set_err(ctx, XT_ERR_CAPACITY,
                     "graph needs " + std::to_string(need) + " bytes of scratch; the allocation-free count holds " +
                         std::to_string(kNoAllocScratchBytes));
    // Every call on this device shares its one array: calls are serialised
    // here, and each returns only after its last kernel has finished.
    std::lock_guard<std::mutex> dl(dev.call_mu);
    Scratch s;
    carve(s, dev.ptr, nv, ne);
    uint64_t* counters = s.counters.as<uint64_t>();
    auto run = [&]() -> int {
      XT_TRY(validate<int64_t>(ctx, nv, ne, d_indices, d_edge_list, false, st, counters));
      XT_TRY(count_device_csr<int64_t>(ctx, s, nv, ne, d_indices, d_edge_list, XT_ORIENT_DEGREE, st, false));
      uint64_t bad = 0;
      XT_HIP(ctx, hipMemcpyAsync(&bad, counters + kBadIds, sizeof(bad), hipMemcpyDeviceToHost, st));
      XT_HIP(ctx, hipStreamSynchronize(st));
      if (bad) return set_err(ctx, XT_ERR_INPUT, "neighbour id out of range [0, nv)");
      return launch_add_total(ctx, d_local_count, st, counters);
    };
    const int rc = run();
    // On every path, nothing queued by this call may still use the array
    // once the device lock is released.
    const hipError_t e = hipStreamSynchronize(st);
    if (rc != XT_OK) return rc;
    if (e != hipSuccess) return hip_err(ctx, e, "hipStreamSynchronize");
    return XT_OK;
  });
}

int xt_single_gpu_count_masked(xt_ctx* ctx, int64_t nedges, const int64_t* d_edge_mask,
                               const int64_t* d_indices, const int64_t* d_edge_list,
                               int64_t* d_local_count, void* stream) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (nedges < 0 || (nedges > 0 && (!d_edge_mask || !d_indices || !d_edge_list)) || !d_local_count)
      return set_err(ctx, XT_ERR_ARG, "bad arguments");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    return launch_masked(ctx, nedges, d_edge_mask, d_indices, d_edge_list, d_local_count, st);
  });
}

int xt_local_count_gpu_inner(xt_ctx* ctx, int64_t start, int64_t end, int64_t jstart, int64_t jend,
                             const int64_t* d_edge_list, int64_t* d_local_count, void* stream) {
  return xt_guarded([&]() -> int {
    if (!ctx) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (!d_edge_list || !d_local_count || start < 0 || jstart < 0) return set_err(ctx, XT_ERR_ARG, "bad arguments");
    // The reference passes the address of a stream (counting.cuh:314-323,
    // xxtric-gpu.hpp:225) and its launcher dereferences it (counting.cu:656).
    hipStream_t st = stream ? *static_cast<hipStream_t*>(stream) : ctx->stream;
    return launch_inner(ctx, start, end, jstart, jend, d_edge_list, d_local_count, st);
  });
}

// ---- resident graphs ----------------------------------------------------
int xt_graph_from_csr(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* offsets, const void* tails,
                      int tails_bytes, uint32_t flags, xt_graph** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    if (tails_bytes != 4 && tails_bytes != 8) return set_err(ctx, XT_ERR_ARG, "tails_bytes must be 4 or 8");
    if (ne > 0 && !tails) return set_err(ctx, XT_ERR_ARG, "null tails");
    XT_TRY(check_host_offsets(ctx, nv, ne, offsets));
    xt_graph* g = new xt_graph();
    g->nv = nv;
    g->ne = ne;
    hipError_t e = hipMalloc(&g->off, (size_t)(nv + 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&g->tails, (size_t)(ne > 0 ? ne : 1) * sizeof(int32_t));
    if (e != hipSuccess) {
      if (g->off) (void)hipFree(g->off);
      delete g;
      return hip_err(ctx, e, "hipMalloc(graph)");
    }
    int rc = upload_csr(ctx, nv, ne, offsets, tails, tails_bytes, g->off, g->tails, ctx->stream);
    if (rc == XT_OK && (flags & XT_CLEAN)) rc = graph_clean(ctx, g);  // may change g->ne
    if (rc == XT_OK && (flags & XT_VALIDATE))
      rc = validate<int32_t>(ctx, g->nv, g->ne, g->off, g->tails, (flags & XT_CHECK_SYMMETRY) != 0, ctx->stream);
    if (rc != XT_OK) {
      (void)hipFree(g->off);
      (void)hipFree(g->tails);
      delete g;
      return rc;
    }
    *out = g;
    return XT_OK;
  });
}

int xt_graph_rmat(xt_ctx* ctx, int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                  int permute, xt_graph** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return set_err(ctx, XT_ERR_ARG, "bad R-MAT probabilities");
    const double s = 4294967296.0;
    const double t[3] = {a * s, (a + b) * s, (a + b + c) * s};
    uint32_t thr[3];
    for (int i = 0; i < 3; i++) thr[i] = t[i] >= s ? 0xFFFFFFFFu : (uint32_t)t[i];
    xt_graph* g = new xt_graph();
    int rc = gen_rmat(ctx, scale, edgefactor, thr, seed, permute, g);
    if (rc != XT_OK) {
      if (g->off) (void)hipFree(g->off);
      if (g->tails) (void)hipFree(g->tails);
      delete g;
      return rc;
    }
    *out = g;
    return XT_OK;
  });
}

int xt_graph_er(xt_ctx* ctx, int64_t n, uint64_t seed, uint64_t thr64, xt_graph** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    xt_graph* g = new xt_graph();
    int rc = gen_er(ctx, n, seed, thr64, g);
    if (rc != XT_OK) {
      if (g->off) (void)hipFree(g->off);
      if (g->tails) (void)hipFree(g->tails);
      delete g;
      return rc;
    }
    *out = g;
    return XT_OK;
  });
}

int xt_graph_from_edges(xt_ctx* ctx, int64_t nv, int64_t m, const int64_t* src, const int64_t* dst,
                        xt_graph** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    if (nv < 1 || nv >= (1LL << 31) - 1 || m < 0 || (m > 0 && (!src || !dst)))
      return set_err(ctx, XT_ERR_ARG, "edge list: need 1 <= nv < 2^31-1, m >= 0 and both id arrays");
    xt_graph* g = new xt_graph();
    const int rc = graph_from_edges(ctx, nv, m, src, dst, g);
    if (rc != XT_OK) {
      if (g->off) (void)hipFree(g->off);
      if (g->tails) (void)hipFree(g->tails);
      delete g;
      return rc;
    }
    *out = g;
    return XT_OK;
  });
}

}  // extern "C"

// vite file (graph.hpp:294-395 layout: E nv, E ne, (nv+1) E offsets rebased to
// 0 as graph.hpp:386-388 does, ne records {E tail, W weight}) streamed into a
// resident graph through two pinned buffers: the next block is read and
// converted to int32 ids while the previous one uploads.
template <typename E, typename W>
static int vite_to_device(xt_ctx* ctx, std::FILE* f, xt_graph* g) {
  hipStream_t st = ctx->stream;
  E hdr[2];
  if (std::fread(hdr, sizeof(E), 2, f) != 2) return set_err(ctx, XT_ERR_INPUT, "vite: short header");
  const int64_t nv = (int64_t)hdr[0], ne = (int64_t)hdr[1];
  if (nv < 1 || nv >= (1LL << 31) - 1 || ne < 0)
    return set_err(ctx, XT_ERR_INPUT, "vite: need 1 <= nv < 2^31-1 and ne >= 0");
  g->nv = nv;
  g->ne = ne;
  XT_HIP(ctx, hipMalloc(&g->off, (size_t)(nv + 1) * sizeof(int64_t)));
  XT_HIP(ctx, hipMalloc(&g->tails, (size_t)(ne > 0 ? ne : 1) * sizeof(int32_t)));
  constexpr int64_t kBlock = 1 << 22;  // elements per pinned block
  struct Rec {
    E tail;
    W w;
  };
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  std::vector<unsigned char> raw((size_t)kBlock * std::max(sizeof(E), sizeof(Rec)));
  int rc = XT_OK;
  auto fail = [&](int code, const char* msg) {
    if (rc == XT_OK) rc = set_err(ctx, code, msg);
  };
  hipError_t e = hipSuccess;
  for (int i = 0; i < 2 && e == hipSuccess; i++) {
    e = hipHostMalloc(&pin[i], (size_t)kBlock * sizeof(int64_t));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) rc = hip_err(ctx, e, "vite staging buffers");
  // Offsets: rebased, checked monotone and ending at ne.
  int64_t base = 0, prev = 0, k = 0;
  for (int64_t i = 0; rc == XT_OK && i <= nv; i += kBlock, k++) {
    const int64_t n = std::min(kBlock, nv + 1 - i);
    const int b = (int)(k & 1);
    if (std::fread(raw.data(), sizeof(E), (size_t)n, f) != (size_t)n) {
      fail(XT_ERR_INPUT, "vite: short offsets");
      break;
    }
    const E* src = reinterpret_cast<const E*>(raw.data());
    if (i == 0) base = prev = (int64_t)src[0];
    if (k >= 2) {
      e = hipEventSynchronize(done[b]);  // this pinned block's previous upload is done
      if (e != hipSuccess) break;
    }
    int64_t* dst = static_cast<int64_t*>(pin[b]);
    int64_t down = 0;
    for (int64_t t = 0; t < n; t++) {
      const int64_t v = (int64_t)src[t];
      down |= (int64_t)(v < prev);
      prev = v;
      dst[t] = v - base;
    }
    if (down) fail(XT_ERR_INPUT, "vite: offsets are not monotone");
    if (i + n == nv + 1 && prev - base != ne) fail(XT_ERR_INPUT, "vite: last offset does not match ne");
    e = hipMemcpyAsync(g->off + i, dst, (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(done[b], st);
    if (e != hipSuccess) break;
  }
  // Ids: int32, checked in [0, nv).
  for (int64_t i = 0; rc == XT_OK && e == hipSuccess && i < ne; i += kBlock, k++) {
    const int64_t n = std::min(kBlock, ne - i);
    const int b = (int)(k & 1);
    if (std::fread(raw.data(), sizeof(Rec), (size_t)n, f) != (size_t)n) {
      fail(XT_ERR_INPUT, "vite: short edge records");
      break;
    }
    const Rec* src = reinterpret_cast<const Rec*>(raw.data());
    if (k >= 2) {
      e = hipEventSynchronize(done[b]);
      if (e != hipSuccess) break;
    }
    int32_t* dst = static_cast<int32_t*>(pin[b]);
    uint32_t out = 0;
    for (int64_t t = 0; t < n; t++) {
      const int64_t v = (int64_t)src[t].tail;
      out |= (uint32_t)(v < 0 || v >= nv);
      dst[t] = (int32_t)v;
    }
    if (out) fail(XT_ERR_INPUT, "vite: neighbour id out of range [0, nv)");
    e = hipMemcpyAsync(g->tails + i, dst, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(done[b], st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess && rc == XT_OK) rc = hip_err(ctx, e, "vite upload");
  for (int i = 0; i < 2; i++) {
    if (done[i]) (void)hipEventDestroy(done[i]);
    if (pin[i]) (void)hipHostFree(pin[i]);
  }
  return rc;
}

extern "C" {

int xt_graph_from_vite(xt_ctx* ctx, const char* path, int layout32, uint32_t flags, xt_graph** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !path || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return set_err(ctx, XT_ERR_ARG, std::string("vite: cannot open ") + path);
    xt_graph* g = new xt_graph();
    int rc = layout32 ? vite_to_device<int32_t, float>(ctx, f, g) : vite_to_device<int64_t, double>(ctx, f, g);
    std::fclose(f);
    if (rc == XT_OK && (flags & XT_CLEAN)) rc = graph_clean(ctx, g);
    if (rc == XT_OK && (flags & XT_VALIDATE))
      rc = validate<int32_t>(ctx, g->nv, g->ne, g->off, g->tails, (flags & XT_CHECK_SYMMETRY) != 0, ctx->stream);
    if (rc != XT_OK) {
      if (g->off) (void)hipFree(g->off);
      if (g->tails) (void)hipFree(g->tails);
      delete g;
      return rc;
    }
    *out = g;
    return XT_OK;
  });
}

int xt_count_edges(xt_ctx* ctx, int64_t nv, int64_t m, const int64_t* src, const int64_t* dst, uint32_t flags,
                   uint64_t* total, uint64_t* per_vertex) {
  return xt_guarded([&]() -> int {
    if (!ctx || !total) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    xt_graph* g = nullptr;
    XT_TRY(xt_graph_from_edges(ctx, nv, m, src, dst, &g));
    const int rc = xt_graph_count(ctx, g, flags, total, per_vertex);
    xt_graph_destroy(ctx, g);
    return rc;
  });
}

int xt_graph_info(const xt_graph* g, int64_t* nv, int64_t* ne) {
  return xt_guarded([&]() -> int {
    if (!g) return XT_ERR_ARG;
    if (nv) *nv = g->nv;
    if (ne) *ne = g->ne;
    return XT_OK;
  });
}

int xt_graph_download(xt_ctx* ctx, const xt_graph* g, int64_t* offsets, int32_t* tails) {
  return xt_guarded([&]() -> int {
    if (!ctx || !g) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (offsets)
      XT_HIP(ctx, hipMemcpy(offsets, g->off, (size_t)(g->nv + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (tails && g->ne > 0)
      XT_HIP(ctx, hipMemcpy(tails, g->tails, (size_t)g->ne * sizeof(int32_t), hipMemcpyDeviceToHost));
    return XT_OK;
  });
}

int xt_graph_count(xt_ctx* ctx, xt_graph* g, uint32_t flags, uint64_t* total, uint64_t* per_vertex) {
  return xt_guarded([&]() -> int {
    if (!ctx || !g || !total) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (per_vertex && !(flags & XT_PER_VERTEX)) return set_err(ctx, XT_ERR_ARG, "per_vertex needs XT_PER_VERTEX");
    hipStream_t st = ctx->stream;
    XT_TRY(count_device_csr<int32_t>(ctx, ctx->s, g->nv, g->ne, g->off, g->tails, flags, st, true));
    return finish_count(ctx, g->nv, flags, st, true, total, per_vertex);
  });
}

int xt_graph_count_owned(xt_ctx* ctx, xt_graph* g, int64_t lo, int64_t hi, uint32_t flags, uint64_t* owned,
                         uint64_t* local, uint64_t* total) {
  return xt_guarded([&]() -> int {
    if (!ctx || !g || !owned) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    if (lo < 0 || hi < lo || hi > g->nv) return set_err(ctx, XT_ERR_ARG, "owned range must satisfy 0 <= lo <= hi <= nv");
    if (flags & XT_PER_VERTEX) return set_err(ctx, XT_ERR_ARG, "an owned-range count has no per-vertex output");
    hipStream_t st = ctx->stream;
    // The owned-range count skips what cannot be owned, so the graph's total
    // takes a count of its own (run first: the timing kept is the owned count's).
    if (total) {
      XT_TRY(count_device_csr<int32_t>(ctx, ctx->s, g->nv, g->ne, g->off, g->tails, flags, st, true));
      XT_TRY(finish_count(ctx, g->nv, flags, st, true, total, nullptr));
    }
    const OwnedRange range{lo, hi};
    XT_TRY(count_device_csr<int32_t>(ctx, ctx->s, g->nv, g->ne, g->off, g->tails, flags, st, true, &range));
    XT_TRY(finish_count(ctx, g->nv, flags, st, true, nullptr, nullptr));
    uint64_t h[2] = {0, 0};
    XT_HIP(ctx, hipMemcpy(h, ctx->s.counters.as<uint64_t>() + kOwned, sizeof(h), hipMemcpyDeviceToHost));
    *owned = h[0];
    if (local) *local = h[1];
    return XT_OK;
  });
}

int xt_graph_slice(xt_ctx* ctx, const xt_graph* g, int64_t lo, int64_t hi, xt_graph** out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !g || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    if (lo < 0 || hi < lo || hi > g->nv) return set_err(ctx, XT_ERR_ARG, "owned range must satisfy 0 <= lo <= hi <= nv");
    xt_graph* s = new xt_graph();
    const int rc = graph_slice(ctx, g, lo, hi, s);
    if (rc != XT_OK) {
      if (s->off) (void)hipFree(s->off);
      if (s->tails) (void)hipFree(s->tails);
      delete s;
      return rc;
    }
    *out = s;
    return XT_OK;
  });
}

int xt_graph_alg_bytes(xt_ctx* ctx, xt_graph* g, uint32_t flags, uint64_t* bytes, uint64_t* sum_min,
                       int64_t* ne_plus) {
  return xt_guarded([&]() -> int {
    if (!ctx || !g) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    XT_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int mode = (flags & XT_ORIENT_ID) ? kOrientId : kOrientDegree;
    XT_TRY(ensure(ctx, ctx->s.counters, kNumSlots * sizeof(uint64_t)));
    XT_TRY(fill_u64(ctx, ctx->s.counters.as<uint64_t>(), kNumSlots, 0, st));
    XT_TRY(orient<int32_t>(ctx, ctx->s, g->nv, g->ne, g->off, g->tails, mode, st));
    XT_TRY(alg_stats(ctx, ctx->s, g->nv, g->ne, st));
    uint64_t sm = 0;
    int64_t nep = 0;
    XT_HIP(ctx, hipMemcpyAsync(&sm, ctx->s.counters.as<uint64_t>() + kStatMin, sizeof(sm),
                               hipMemcpyDeviceToHost, st));
    if (g->nv > 0)
      XT_HIP(ctx, hipMemcpyAsync(&nep, ctx->s.ooff.as<int64_t>() + g->nv, sizeof(nep), hipMemcpyDeviceToHost, st));
    XT_HIP(ctx, hipStreamSynchronize(st));
    if (sum_min) *sum_min = sm;
    if (ne_plus) *ne_plus = nep;
    if (bytes) *bytes = 8ULL * (uint64_t)(g->nv + 1) + 4ULL * (uint64_t)nep + 4ULL * sm;
    return XT_OK;
  });
}

int xt_graph_stats_get(xt_ctx* ctx, xt_graph* g, uint32_t flags, xt_graph_stats* out) {
  return xt_guarded([&]() -> int {
    if (!ctx || !g || !out) return XT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    xt_graph_stats st = {};
    XT_TRY(xt_graph_alg_bytes(ctx, g, flags, &st.alg_bytes, &st.sum_min, &st.ne_plus));
    uint64_t h[2] = {0, 0};
    XT_HIP(ctx, hipMemcpy(h, ctx->s.counters.as<uint64_t>() + kStatMinHeavy, sizeof(h),
                          hipMemcpyDeviceToHost));
    st.sum_min_heavy = h[0];
    st.heavy_edges = (int64_t)h[1];
    st.group_bytes = 4ULL * (st.sum_min_heavy + (uint64_t)st.heavy_edges);
    *out = st;
    return XT_OK;
  });
}

int xt_graph_destroy(xt_ctx* ctx, xt_graph* g) {
  return xt_guarded([&]() -> int {
    if (!g) return XT_ERR_ARG;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    if (g->off) (void)hipFree(g->off);
    if (g->tails) (void)hipFree(g->tails);
    delete g;
    return XT_OK;
  });
}

}  // extern "C"
