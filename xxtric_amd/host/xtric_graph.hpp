// xtric_graph.hpp -- single-process host mirror of the reference's graph ingest
// (graph.hpp:19-1374): Graph (CSR of Edge{tail, weight}), BinaryEdgeList (vite
// reader) and GenerateRGG (seeded random geometric graph), built on the
// library's C ABI (include/xtric.h).  Same member names and argument meaning,
// without MPI: one process owns the whole graph (nprocs = 1, rank 0).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "xtric.h"

using GraphElem = int64_t;     // deftypes.hpp:47
using GraphWeight = double;    // deftypes.hpp:48

struct Edge {                  // graph.hpp:20-25
  GraphElem tail_;
  GraphWeight weight_;
  Edge() : tail_(-1), weight_(0.0) {}
};

class Graph {                  // graph.hpp:52-275, one process
 public:
  Graph(GraphElem nv, GraphElem ne) : nv_(nv), ne_(ne) {
    edge_indices_.assign(nv + 1, 0);
    edge_list_.resize(ne);
  }
  void edge_range(GraphElem const vertex, GraphElem& e0, GraphElem& e1) const {
    e0 = edge_indices_[vertex];
    e1 = edge_indices_[vertex + 1];
  }
  Edge const& get_edge(GraphElem const index) const { return edge_list_[index]; }
  Edge& set_edge(GraphElem const index) { return edge_list_[index]; }
  GraphElem get_lnv() const { return nv_; }
  GraphElem get_lne() const { return ne_; }
  GraphElem get_nv() const { return nv_; }
  GraphElem get_ne() const { return ne_; }
  GraphElem get_base(int) const { return 0; }
  GraphElem local_to_global(GraphElem idx) const { return idx; }
  GraphElem global_to_local(GraphElem idx) const { return idx; }
  int get_owner(GraphElem) const { return 0; }

  // Same report as graph.hpp:230-264 at one process.
  void print_dist_stats() const {
    const double average = (double)ne_;
    const double avg_sq = (double)ne_ * (double)ne_;
    const double var = std::abs(avg_sq - average * average);
    std::cout << std::endl;
    std::cout << "-------------------------------------------------------" << std::endl;
    std::cout << "Graph edge distribution characteristics" << std::endl;
    std::cout << "-------------------------------------------------------" << std::endl;
    std::cout << "Number of vertices: " << nv_ << std::endl;
    std::cout << "Number of edges: " << ne_ << std::endl;
    std::cout << "Maximum number of edges: " << ne_ << std::endl;
    std::cout << "Average number of edges: " << average << std::endl;
    std::cout << "Expected value of X^2: " << avg_sq << std::endl;
    std::cout << "Variance: " << var << std::endl;
    std::cout << "Standard deviation: " << std::sqrt(var) << std::endl;
    std::cout << "-------------------------------------------------------" << std::endl;
  }

  // Plain arrays for the C ABI.
  std::vector<int64_t> tails() const {
    std::vector<int64_t> t(ne_);
    for (GraphElem e = 0; e < ne_; e++) t[e] = edge_list_[e].tail_;
    return t;
  }

  std::vector<GraphElem> edge_indices_;
  std::vector<Edge> edge_list_;

 private:
  GraphElem nv_, ne_;
};

inline Graph* graph_from_csr(const std::vector<int64_t>& off, const std::vector<int64_t>& tails) {
  const GraphElem nv = (GraphElem)off.size() - 1;
  Graph* g = new Graph(nv, (GraphElem)tails.size());
  g->edge_indices_ = off;
  for (size_t e = 0; e < tails.size(); e++) {
    g->edge_list_[e].tail_ = tails[e];
    g->edge_list_[e].weight_ = 1.0;
  }
  return g;
}

class BinaryEdgeList {         // graph.hpp:279-395 (read), one process
 public:
  Graph* read(int /*me*/, int /*nprocs*/, int /*ranks_per_node*/, std::string file,
              bool layout32 = false) {
    int64_t nv = 0, ne = 0;
    if (xt_vite_read(file.c_str(), layout32, &nv, &ne, nullptr, nullptr) != XT_OK)
      throw std::runtime_error("Error opening file! " + file);
    std::vector<int64_t> off(nv + 1), tails(ne);
    if (xt_vite_read(file.c_str(), layout32, &nv, &ne, off.data(), tails.data()) != XT_OK)
      throw std::runtime_error("Error reading file! " + file);
    return graph_from_csr(off, tails);
  }

  // Owned vertex bins of a balanced read over nprocs ranks (mbins: nprocs + 1
  // entries; rank r owns [mbins[r], mbins[r+1])): the greedy fill of flag -b
  // (graph.hpp:556-602) and the linear partitioning of flag -d
  // (graph.hpp:397-419).  Offsets are read rebased to 0, as every file the
  // reference writes has them.
  void find_balanced_num_edges(int nprocs, std::string file, std::vector<GraphElem>& mbins,
                               bool layout32 = false) {
    bins_of(nprocs, file, mbins, XT_PARTITION_GREEDY, layout32);
  }
  void linear_partitioning_num_edges(int nprocs, std::string file, std::vector<GraphElem>& mbins,
                                     bool layout32 = false) {
    bins_of(nprocs, file, mbins, XT_PARTITION_LINEAR, layout32);
  }

  // graph.hpp:608-733 at one process (nprocs = 1): rank 0 owns every vertex,
  // so the balanced read is the plain read.
  Graph* read_balanced(int me, int nprocs, int ranks_per_node, std::string file,
                       short /*static_or_dyn*/ = 0, bool layout32 = false) {
    if (nprocs != 1) throw std::runtime_error("read_balanced: one process (nprocs = 1)");
    return read(me, nprocs, ranks_per_node, file, layout32);
  }

 private:
  void bins_of(int nprocs, const std::string& file, std::vector<GraphElem>& mbins, int method,
               bool layout32) {
    int64_t nv = 0, ne = 0;
    if (xt_vite_read(file.c_str(), layout32, &nv, &ne, nullptr, nullptr) != XT_OK)
      throw std::runtime_error("Error opening file! " + file);
    std::vector<int64_t> off(nv + 1);
    if (xt_vite_read(file.c_str(), layout32, &nv, &ne, off.data(), nullptr) != XT_OK)
      throw std::runtime_error("Error reading file! " + file);
    mbins.assign(nprocs + 1, 0);
    const int rc = xt_partition_vertices(nv, off.data(), nprocs, method, mbins.data());
    if (rc != XT_OK) throw std::runtime_error("vertex partition failed (" + std::to_string(rc) + ")");
  }
};

class GenerateRGG {            // graph.hpp:745-1374
 public:
  // nprocs > 1: the graph the reference generates on that many MPI ranks
  // (its constructor's checks, graph.hpp:765-793), all ranks' rows here.
  // rgg_flags: XT_RGG_STRICT (distance < r) and XT_RGG_CONNECT (r = sqrt(2 ln n /
  // (pi n)) when radius <= 0); 0 is the reference's generator.
  explicit GenerateRGG(GraphElem nv, double radius = 0.0, int nprocs = 1, uint32_t rgg_flags = 0)
      : nv_(nv), radius_(radius), nprocs_(nprocs), rgg_flags_(rgg_flags) {
    rn_ = radius > 0 ? radius : (rgg_flags & XT_RGG_CONNECT) ? xt_rgg_radius_connect(nv) : xt_rgg_radius(nv);
    if (nprocs_ < 1 || (nprocs_ & (nprocs_ - 1)) != 0)
      throw std::runtime_error("[ERROR] Number of processes must be a power of 2.");
    if (nv_ % nprocs_ != 0)
      throw std::runtime_error("[ERROR] Number of vertices must be perfectly divisible by number of processes.");
    if (!((1.0 / (double)nprocs_) > rn_))
      throw std::runtime_error("[ERROR] The radius must be below 1/(number of processes).");
  }
  // isLCG: the reference's seeded LCG stream (x0 = seed_seq{1}); otherwise a
  // distinct seeded stream (seed 2) -- the reference draws from random_device
  // there and is not reproducible.  randomEdgePercent adds seeded random edges
  // between distinct, not yet adjacent vertex pairs (graph.hpp:1099-1283 adds
  // them from a time/pid seed).
  Graph* generate(bool isLCG, bool /*unitEdgeWeight*/ = true, GraphWeight randomEdgePercent = 0.0,
                  uint32_t seed = 0) {
    const uint32_t s = seed ? seed : (isLCG ? 1u : 2u);
    int64_t ne = 0;
    std::vector<int64_t> off(nv_ + 1);
    if (xt_rgg_generate_opts(nv_, nprocs_, s, radius_, rgg_flags_, off.data(), nullptr, 0, &ne) != XT_OK)
      throw std::runtime_error("RGG generation failed");
    std::vector<int64_t> tails(ne);
    xt_rgg_generate_opts(nv_, nprocs_, s, radius_, rgg_flags_, off.data(), tails.data(), ne, &ne);
    if (randomEdgePercent > 0.0) add_random_edges(off, tails, randomEdgePercent, s);
    return graph_from_csr(off, tails);
  }
  GraphWeight get_d() const { return rn_; }
  GraphElem get_nv() const { return nv_; }

 private:
  void add_random_edges(std::vector<int64_t>& off, std::vector<int64_t>& tails, double pct,
                        uint32_t seed) {
    std::vector<std::set<int64_t>> rows(nv_);
    for (GraphElem i = 0; i < nv_; i++) rows[i].insert(tails.begin() + off[i], tails.begin() + off[i + 1]);
    const GraphElem nrande = (GraphElem)(pct * (double)(tails.size() / 2)) / 100;
    std::mt19937_64 re(0x5eed0000ULL ^ seed);
    std::uniform_int_distribution<GraphElem> R(0, nv_ - 1);
    GraphElem added = 0;
    for (GraphElem k = 0; k < nrande; k++) {
      const GraphElem i = R(re), j = R(re);
      if (i == j || rows[i].count(j)) continue;
      rows[i].insert(j);
      rows[j].insert(i);
      added++;
    }
    std::cout << "Adding extra " << added << " edges while trying to incorporate " << pct << "%"
              << " extra edges globally." << std::endl;
    tails.clear();
    off[0] = 0;
    for (GraphElem i = 0; i < nv_; i++) {
      tails.insert(tails.end(), rows[i].begin(), rows[i].end());
      off[i + 1] = (int64_t)tails.size();
    }
  }
  GraphElem nv_;
  double radius_, rn_;
  int nprocs_;
  uint32_t rgg_flags_;
};
