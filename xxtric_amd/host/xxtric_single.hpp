// xxtric_single.hpp -- drop-in for the reference's single-device engine
// TriangleCountingSingle (xxtric-gpu.hpp:22-298) over the C ABI.
//
//   TriangleCountingSingle tr(g);   // uploads the CSR to HBM (xxtric-gpu.hpp:27-156)
//   GraphElem n = tr.count();       // device count, synchronised (xxtric-gpu.hpp:192-270)
//   tr.clear();                     // frees device memory (xxtric-gpu.hpp:272-285)
//
// The engine reads the graph only through the members the reference's Graph
// (graph.hpp:52-275) makes public -- get_nv(), get_ne(), edge_indices_ and
// get_edge(e).tail_ -- so it takes that class unchanged: main.cpp includes
// graph.hpp (main.cpp:13-15) and then this header in place of xxtric-gpu.hpp.
// This header declares no graph type of its own; the one-process mirror in
// xtric_graph.hpp (used by the tric driver) has the same members and works
// too.  GraphElem may be int32 (USE_32_BIT_GRAPH, deftypes.hpp:40-45) or
// int64: offsets and ids are widened to int64 for the C ABI.
//
// Orientation and intersection run on the MI355X; count() returns the exact
// global count.  per_vertex() adds the per-vertex counts with the reference's
// smallest-vertex credit (xxtric.hpp:160-206), which the reference keeps
// internal.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "xtric.h"

class TriangleCountingSingle {
 public:
  template <class G>
  explicit TriangleCountingSingle(G* g, uint32_t flags = XT_ORIENT_DEGREE, int device = 0)
      : flags_(flags) {
    if (!g) throw std::invalid_argument("TriangleCountingSingle: null graph");
    const int64_t nv = (int64_t)g->get_nv(), ne = (int64_t)g->get_ne();
    if ((int64_t)g->edge_indices_.size() != nv + 1)
      throw std::invalid_argument("TriangleCountingSingle: edge_indices_ must hold nv + 1 offsets");
    std::vector<int64_t> off(g->edge_indices_.begin(), g->edge_indices_.end());
    std::vector<int64_t> tails((size_t)ne);
    for (int64_t e = 0; e < ne; e++) tails[(size_t)e] = (int64_t)g->get_edge(e).tail_;
    // A constructor that throws runs no destructor: release the context here.
    try {
      check(xt_ctx_create(device, &ctx_), "xt_ctx_create");
      check(xt_graph_from_csr(ctx_, nv, ne, off.data(), tails.data(), 8, XT_VALIDATE, &dg_),
            "xt_graph_from_csr");
    } catch (...) {
      clear();
      throw;
    }
    nv_ = nv;
  }
  // Adopts a graph already resident on the device (e.g. xt_graph_rmat); the
  // caller keeps the context.
  TriangleCountingSingle(xt_ctx* ctx, xt_graph* dg, uint32_t flags = XT_ORIENT_DEGREE)
      : ctx_(ctx), dg_(dg), owns_ctx_(false), flags_(flags) {
    int64_t ne = 0;
    xt_graph_info(dg, &nv_, &ne);
  }
  TriangleCountingSingle(const TriangleCountingSingle&) = delete;
  TriangleCountingSingle& operator=(const TriangleCountingSingle&) = delete;
  ~TriangleCountingSingle() { clear(); }

  // The reference returns GraphElem; int64_t converts to it either way.
  int64_t count() {
    uint64_t total = 0;
    check(xt_graph_count(ctx_, dg_, flags_ & ~XT_PER_VERTEX, &total, nullptr), "xt_graph_count");
    return (int64_t)total;
  }

  std::vector<uint64_t> per_vertex() {
    std::vector<uint64_t> pv((size_t)nv_);
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
