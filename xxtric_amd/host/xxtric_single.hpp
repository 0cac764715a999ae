// xxtric_single.hpp -- drop-in for the reference's single-device engine
// TriangleCountingSingle (xxtric-gpu.hpp:22-298) over the C ABI.
//
//   TriangleCountingSingle tr(g);   // uploads the CSR to HBM (xxtric-gpu.hpp:27-156)
//   GraphElem n = tr.count();       // device count, synchronised (xxtric-gpu.hpp:192-270)
//   tr.clear();                     // frees device memory (xxtric-gpu.hpp:272-285)
//
// Orientation and intersection run on the MI355X; count() returns the exact
// global count.  per_vertex() adds the per-vertex counts with the reference's
// smallest-vertex credit (xxtric.hpp:160-206), which the reference keeps
// internal.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "xtric.h"
#include "xtric_graph.hpp"

class TriangleCountingSingle {
 public:
  explicit TriangleCountingSingle(Graph* g, uint32_t flags = XT_ORIENT_DEGREE, int device = 0)
      : flags_(flags) {
    check(xt_ctx_create(device, &ctx_), "xt_ctx_create");
    const std::vector<int64_t> t = g->tails();
    check(xt_graph_from_csr(ctx_, g->get_nv(), g->get_ne(), g->edge_indices_.data(), t.data(), 8,
                            XT_VALIDATE, &dg_),
          "xt_graph_from_csr");
    nv_ = g->get_nv();
  }
  // Adopts a graph already resident on the device (e.g. xt_graph_rmat).
  TriangleCountingSingle(xt_ctx* ctx, xt_graph* dg, uint32_t flags = XT_ORIENT_DEGREE)
      : ctx_(ctx), dg_(dg), owns_ctx_(false), flags_(flags) {
    int64_t ne = 0;
    xt_graph_info(dg, &nv_, &ne);
  }
  ~TriangleCountingSingle() { clear(); }

  GraphElem count() {
    uint64_t total = 0;
    check(xt_graph_count(ctx_, dg_, flags_ & ~XT_PER_VERTEX, &total, nullptr), "xt_graph_count");
    return (GraphElem)total;
  }

  std::vector<uint64_t> per_vertex() {
    std::vector<uint64_t> pv(nv_);
    uint64_t total = 0;
    check(xt_graph_count(ctx_, dg_, flags_ | XT_PER_VERTEX, &total, pv.data()), "xt_graph_count");
    return pv;
  }

  xt_timing timing() const {
    xt_timing t{};
    xt_last_timing(ctx_, &t);
    return t;
  }

  void clear() {
    if (dg_) xt_graph_destroy(ctx_, dg_);
    dg_ = nullptr;
    if (ctx_ && owns_ctx_) xt_ctx_destroy(ctx_);
    ctx_ = nullptr;
  }

 private:
  void check(int rc, const char* what) {
    if (rc != XT_OK)
      throw std::runtime_error(std::string(what) + ": " + (ctx_ ? xt_last_error(ctx_) : "no device"));
  }
  xt_ctx* ctx_ = nullptr;
  xt_graph* dg_ = nullptr;
  bool owns_ctx_ = true;
  uint32_t flags_ = XT_ORIENT_DEGREE;
  int64_t nv_ = 0;
};
