// counting_xtric.hpp -- the three host launchers of the reference's
// counting.cuh (single_gpu_count :259-265, single_gpu_count_masked :229-245,
// local_count_gpu_inner :314-323) on top of libxtric's C ABI, for a
// maintainer who keeps the reference engine's own call sites
// (xxtric-gpu.hpp:195-263) and replaces only the nvcc-built counting.o.
//
// Same names, argument order and meaning.  Differences:
//   * streams are HIP streams: the masked launcher takes one by value, the
//     inner launcher the ADDRESS of one, as the reference's cudaStream_t*
//     (its call site passes &streams[id % MAX_CUDA_STREAMS]); both are
//     templates, so this header needs no HIP include;
//   * a failed call prints the library's message and exits with its status
//     (the reference checks nothing unless built with -DDEBUG_CUDA,
//     counting.cuh:9-23, which then exits the same way);
//   * the library context is created once, on first use, by a function-local
//     static (thread-safe initialisation): the reference's USE_SPLIT_KERNELS
//     loop calls the inner launcher from an OpenMP parallel loop.  Calls on
//     the context are serialised by the library.
// GraphElem must be int64 here (the reference's default, deftypes.hpp:47):
// the launchers take the device arrays as they are.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "xtric.h"

namespace xtric_shim {

inline xt_ctx* context() {
  static xt_ctx* const ctx = [] {
    xt_ctx* c = nullptr;
    const int rc = xt_ctx_create(0, &c);
    if (rc != XT_OK) {
      std::fprintf(stderr, "libxtric: xt_ctx_create(0) failed (%d)\n", rc);
      std::exit(rc < 0 ? -rc : rc);
    }
    return c;
  }();
  return ctx;
}

inline void check(int rc, const char* what) {
  if (rc != XT_OK) {
    std::fprintf(stderr, "libxtric: %s failed (%d): %s\n", what, rc, xt_last_error(context()));
    std::exit(rc < 0 ? -rc : rc);
  }
}

}  // namespace xtric_shim

static_assert(sizeof(GraphElem) == sizeof(int64_t), "the launchers take int64 GraphElem arrays");

inline void single_gpu_count(const GraphElem nv, const GraphElem* d_indices, const GraphElem* d_edge_list,
                             GraphElem* d_local_count) {
  xtric_shim::check(xt_single_gpu_count(xtric_shim::context(), nv, (const int64_t*)d_indices,
                                        (const int64_t*)d_edge_list, (int64_t*)d_local_count, nullptr),
                    "single_gpu_count");
}

inline void single_gpu_count_masked(const GraphElem nedges, const GraphElem* d_edge_mask,
                                    const GraphElem* d_indices, const GraphElem* d_edge_list,
                                    GraphElem* d_local_count) {
  xtric_shim::check(xt_single_gpu_count_masked(xtric_shim::context(), nedges, (const int64_t*)d_edge_mask,
                                               (const int64_t*)d_indices, (const int64_t*)d_edge_list,
                                               (int64_t*)d_local_count, nullptr),
                    "single_gpu_count_masked");
}

template <class Stream>
inline void single_gpu_count_masked(const GraphElem nedges, const GraphElem* d_edge_mask,
                                    const GraphElem* d_indices, const GraphElem* d_edge_list,
                                    GraphElem* d_local_count, Stream stream) {
  xtric_shim::check(xt_single_gpu_count_masked(xtric_shim::context(), nedges, (const int64_t*)d_edge_mask,
                                               (const int64_t*)d_indices, (const int64_t*)d_edge_list,
                                               (int64_t*)d_local_count, (void*)stream),
                    "single_gpu_count_masked");
}

template <class Stream>
inline void local_count_gpu_inner(const GraphElem start, const GraphElem end, const GraphElem jstart,
                                  const GraphElem jend, const GraphElem* d_edge_list, GraphElem* d_local_count,
                                  Stream* stream) {
  xtric_shim::check(xt_local_count_gpu_inner(xtric_shim::context(), start, end, jstart, jend,
                                             (const int64_t*)d_edge_list, (int64_t*)d_local_count,
                                             (void*)stream),
                    "local_count_gpu_inner");
}
