// xt_device.hpp -- device-side helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xt {

constexpr int kWave = 64;  // CDNA wavefront width

// XCD-aware block order.  The chip's eight XCDs (each with its own L2) receive
// workgroups round-robin by workgroup index, so logical block
// (b mod 8) * per + b / 8 hands XCD x the contiguous blocks [x per, (x+1) per):
// neighbouring chunks of the input -- rows of the same small graph in a
// batch -- are read through one L2 instead of eight.  Launch xcd_grid(n)
// workgroups; a logical block >= n does nothing.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks) {
  const int64_t b = blockIdx.x;
  const int64_t per = (nblocks + kXcds - 1) / kXcds;
  return (b % kXcds) * per + b / kXcds;
}
inline unsigned xcd_grid(int64_t nblocks) {
  return (unsigned)(((nblocks + kXcds - 1) / kXcds) * kXcds);
}

// Index of the row containing entry e: the largest r in [lo, hi] with
// off[r] <= e (rows may be empty, so equal offsets are skipped to the last).
__device__ __forceinline__ int64_t row_of(const int64_t* __restrict__ off, int64_t lo, int64_t hi,
                                          int64_t e) {
  // invariant: off[lo] <= e, answer in [lo, hi]
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo + 1) >> 1);
    if (off[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// First and last row of input chunk c, read from the per-chunk table that
// k_chunk_rows builds (clamped to [0, nv), so a table built from an invalid
// index still keeps every access in range).
struct RowSpan {
  int64_t lo, hi;
};
__device__ __forceinline__ RowSpan chunk_span(const int64_t* __restrict__ crows, int64_t c, int64_t nv) {
  int64_t lo = crows[2 * c], hi = crows[2 * c + 1];
  lo = lo < 0 ? 0 : (lo > nv - 1 ? nv - 1 : lo);
  hi = hi < lo ? lo : (hi > nv - 1 ? nv - 1 : hi);
  return RowSpan{lo, hi};
}

// Same search over an LDS copy of off[base .. base+n] (n+1 entries).
__device__ __forceinline__ int lds_row_of(const int64_t* s_off, int n, int64_t e) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = lo + ((hi - lo + 1) >> 1);
    if (s_off[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// First index in [lo, hi) with a[idx] >= x (hi if none).
template <typename T>
__device__ __forceinline__ int64_t lower_bound_dev(const T* __restrict__ a, int64_t lo, int64_t hi,
                                                   T x) {
  int64_t len = hi - lo;
  while (len > 0) {
    const int64_t half = len >> 1;
    if (a[lo + half] < x) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

// Block-wide sum (blockDim.x threads, multiple of 64); result valid in thread 0.
template <int kThreads>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* s_tmp /* kThreads/64 */) {
  v = wave_sum_u64(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) s_tmp[wid] = v;
  __syncthreads();
  uint64_t r = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kThreads / kWave; i++) r += s_tmp[i];
  }
  return r;
}

// Inclusive scan across a wave of 64-bit values.
__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int64_t y = __shfl_up(v, off, kWave);
    if (lane >= off) v += y;
  }
  return v;
}

// Exclusive block scan of one value per thread; *total receives the block sum.
template <int kThreads>
__device__ __forceinline__ int64_t block_excl_scan_i64(int64_t v, int64_t* s_tmp /* kThreads/64 */,
                                                       int64_t* total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int64_t inc = wave_incl_scan_i64(v);
  if (lane == kWave - 1) s_tmp[wid] = inc;
  __syncthreads();
  int64_t before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kThreads / kWave; i++) {
    const int64_t t = s_tmp[i];
    if (i < wid) before += t;
    all += t;
  }
  *total = all;
  return before + inc - v;
}

// Loads the kPerThread (8) entries of one thread; a full aligned run of int32
// ids is fetched as two 16-byte loads.
template <typename T, int kRun>
__device__ __forceinline__ void load_run(const T* __restrict__ tails, int64_t e0, int64_t ee,
                                         int64_t (&vals)[kRun]) {
  static_assert(kRun == 8, "runs of eight entries");
  if constexpr (sizeof(T) == 4) {
    if (e0 + kRun <= ee) {
      const int4* p = reinterpret_cast<const int4*>(tails + e0);
      const int4 a = p[0], b = p[1];
      vals[0] = a.x; vals[1] = a.y; vals[2] = a.z; vals[3] = a.w;
      vals[4] = b.x; vals[5] = b.y; vals[6] = b.z; vals[7] = b.w;
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < kRun; j++) vals[j] = (e0 + j < ee) ? (int64_t)tails[e0 + j] : 0;
}

// SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014).
__device__ __forceinline__ uint64_t mix64_dev(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace xt
