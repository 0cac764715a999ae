// tric.cpp -- command-line driver, flag- and output-compatible with the
// reference's main.cpp (single-device branch, main.cpp:49-276), one process.
//
// Reference flags (main.cpp:204-236):
//   -f FILE   read a vite binary edge list          -n N   generate an RGG of N vertices
//   -l        use the seeded LCG stream (reference)  -p P   add P% random edges (seeded here)
//   -b / -d   balanced reads: greedy / linear-partitioning vertex bins (one process:
//             the read itself is a plain read)
//   -r R      ranks per node (ignored: one process)  -s B   buffer size (ignored)
// Additional flags of this build:
//   -R S      R-MAT scale S generated on the device (with -e edge factor, -S seed, -P no relabel)
//   -m MODE   orientation: degree (default) or id   -g D   HIP device ordinal
//   -v FILE   write per-vertex counts (one line per vertex)
//   -w FILE   write the input graph as a vite file  -x     generate/read only, no count
//   -L        read/write the 32-bit vite layout      -j     also print a JSON summary line
//   -t K      time K counts (the reference times one)
//   -B K      with -f: print the vertex bins of a balanced read over K ranks
//             (greedy, or linear partitioning with -d)
//   -N P      with -n: generate the graph the reference builds on P MPI ranks
//             (`mpiexec -n P`: per-rank LCG streams and strips, cross-rank
//             edges as its exchange leaves them), all of it counted here
//   -C        with -n: radius sqrt(2 ln N / (pi N)) and edges at distance < r
//             (the reference uses graph.hpp:789-791 and <=)
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "xtric_graph.hpp"
#include "xxtric_single.hpp"

namespace {

struct Opts {
  std::string file, vfile, wfile;
  GraphElem nv = 0;
  bool lcg = false, layout32 = false, json = false, no_count = false, permute = true;
  double pct = 0.0;
  int rmat = 0, ef = 16, device = 0, reps = 1;
  int balanced = 0;   // 1: -b (greedy bins), 2: -d (linear partitioning)
  int bin_ranks = 0;  // -B K
  int gen_ranks = 1;  // -N P
  uint32_t rgg_flags = 0;  // -C
  uint64_t seed = 1;
  bool seed_set = false;
  uint32_t flags = XT_ORIENT_DEGREE;
};

[[noreturn]] void die(const std::string& m) {
  std::cerr << m << std::endl;
  std::exit(1);
}

Opts parse(int argc, char* argv[]) {
  Opts o;
  int c;
  while ((c = getopt(argc, argv, "f:r:n:p:olbds:R:e:S:Pm:g:v:w:xLjt:B:N:C")) != -1) {
    switch (c) {
      case 'f': o.file = optarg; break;
      case 'b': o.balanced = 1; break;
      case 'd': o.balanced = 2; break;
      case 'o': break;
      case 'B': o.bin_ranks = atoi(optarg); break;
      case 'N': o.gen_ranks = atoi(optarg); break;
      case 'C': o.rgg_flags = XT_RGG_STRICT | XT_RGG_CONNECT; break;
      case 'r': case 's': break;             // ranks per node / buffer size: no meaning at one process
      case 'n': o.nv = atol(optarg); break;
      case 'l': o.lcg = true; break;
      case 'p': o.pct = atof(optarg); break;
      case 'R': o.rmat = atoi(optarg); break;
      case 'e': o.ef = atoi(optarg); break;
      case 'S': o.seed = strtoull(optarg, nullptr, 10); o.seed_set = true; break;
      case 'P': o.permute = false; break;
      case 'm': o.flags = std::string(optarg) == "id" ? XT_ORIENT_ID : XT_ORIENT_DEGREE; break;
      case 'g': o.device = atoi(optarg); break;
      case 'v': o.vfile = optarg; break;
      case 'w': o.wfile = optarg; break;
      case 'x': o.no_count = true; break;
      case 'L': o.layout32 = true; break;
      case 'j': o.json = true; break;
      case 't': o.reps = atoi(optarg) > 0 ? atoi(optarg) : 1; break;
      default: die("Unknown option");
    }
  }
  // main.cpp:255-275
  if (argc == 1) die("Must specify some options.");
  if (o.nv <= 0 && o.file.empty() && o.rmat == 0)
    die("Must specify a binary file name with -f or provide parameters for generating a graph.");
  if (o.pct < 0.0 || o.pct >= 100.0) die("Invalid random edge percentage for generated graph!");
  if (o.pct > 0.0 && o.nv <= 0) die("Must specify -n for graph generation first to add random edges to it.");
  return o;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char* argv[]) {
  Opts o = parse(argc, argv);
  Graph* g = nullptr;
  xt_ctx* ctx = nullptr;
  xt_graph* dg = nullptr;
  const double td0 = now_s();
  try {
    if (o.rmat > 0) {
      if (xt_ctx_create(o.device, &ctx) != XT_OK) die("no HIP device");
      if (xt_graph_rmat(ctx, o.rmat, o.ef, 0.57, 0.19, 0.19, o.seed, o.permute, &dg) != XT_OK)
        die(std::string("R-MAT generation failed: ") + xt_last_error(ctx));
    } else if (o.nv > 0) {
      GenerateRGG gr(o.nv, 0.0, o.gen_ranks, o.rgg_flags);
      g = gr.generate(o.lcg, true, o.pct, o.seed_set ? (uint32_t)o.seed : 0);
    } else {
      BinaryEdgeList rm;
      if (o.balanced) {  // main.cpp:95-103 and graph.hpp:620-636, rank 0 of one process
        std::cout << std::endl;
        std::cout << "Trying to balance the edge distribution while reading: " << std::endl;
        std::cout << o.file << std::endl;
        if (o.balanced == 1)
          std::cout << "Trying to achieve equal edge distribution across processes." << std::endl;
        else
          std::cout << "Trying to achieve equal edge distribution across processes via dynamic "
                       "programming (linear partitioning)." << std::endl;
        g = rm.read_balanced(0, 1, 1, o.file, (short)(o.balanced - 1), o.layout32);
      } else {
        g = rm.read(0, 1, 1, o.file, o.layout32);
      }
      if (o.bin_ranks > 0) {
        std::vector<GraphElem> bins;
        if (o.balanced == 2)
          rm.linear_partitioning_num_edges(o.bin_ranks, o.file, bins, o.layout32);
        else
          rm.find_balanced_num_edges(o.bin_ranks, o.file, bins, o.layout32);
        std::cout << "Vertex bins for " << o.bin_ranks << " ranks:";
        for (GraphElem b : bins) std::cout << " " << b;
        std::cout << std::endl;
      }
    }
  } catch (const std::exception& e) {
    die(e.what());
  }
  const double td = now_s() - td0;
  int64_t nv = 0, ne = 0;
  if (g) {
    g->print_dist_stats();
    nv = g->get_nv();
    ne = g->get_ne();
  } else {
    xt_graph_info(dg, &nv, &ne);
    Graph stats(nv, ne);
    stats.print_dist_stats();
  }
  if (o.file.empty())
    std::cout << "Time to generate distributed graph of " << nv << " vertices (secs.): " << td << std::endl;
  else
    std::cout << "Time to read input file and create distributed graph (secs.): " << td << std::endl;

  if (!o.wfile.empty()) {
    std::vector<int64_t> off, tails;
    if (g) {
      off = g->edge_indices_;
      tails = g->tails();
    } else {
      off.resize(nv + 1);
      std::vector<int32_t> t32(ne);
      xt_graph_download(ctx, dg, off.data(), t32.data());
      tails.assign(t32.begin(), t32.end());
    }
    if (xt_vite_write(o.wfile.c_str(), o.layout32, nv, off.data(), tails.data()) != XT_OK)
      die("could not write " + o.wfile);
  }
  if (o.no_count) return 0;

  try {
    TriangleCountingSingle* tr = g ? new TriangleCountingSingle(g, o.flags, o.device)
                                   : new TriangleCountingSingle(ctx, dg, o.flags);
    GraphElem ntris = 0;
    double tmin = 1e300, tmax = 0, tsum = 0;
    for (int r = 0; r < o.reps; r++) {
      const double t0 = now_s();
      ntris = tr->count();  // main.cpp:160-162
      const double t = now_s() - t0;
      tmin = std::min(tmin, t);
      tmax = std::max(tmax, t);
      tsum += t;
    }
    const double avg_t = tsum / o.reps;
    std::cout << "Avg/Min/Max execution time (secs.) for triangle counting on 1 processes: " << avg_t
              << ", " << tmin << ", " << tmax << std::endl;
    std::cout << "Number of triangles: " << ntris << std::endl;
    std::cout << "TEPS: " << (double)ne / avg_t << std::endl;   // main.cpp:188 (ne = directed entries)
    std::cout << "Resolution of timer: 1e-09" << std::endl;
    const xt_timing t = tr->timing();
    if (!o.vfile.empty()) {
      std::vector<uint64_t> pv = tr->per_vertex();
      std::ofstream f(o.vfile);
      for (size_t i = 0; i < pv.size(); i++) f << pv[i] << "\n";
    }
    if (o.json)
      std::cout << "{\"triangles\": " << ntris << ", \"nv\": " << nv << ", \"ne\": " << ne
                << ", \"count_s\": " << avg_t << ", \"orient_ms\": " << t.orient_ms
                << ", \"intersect_ms\": " << t.count_ms << ", \"device_ms\": " << t.total_ms
                << ", \"ne_plus\": " << t.ne_plus << "}" << std::endl;
    tr->clear();
    delete tr;
  } catch (const std::exception& e) {
    die(e.what());
  }
  delete g;
  if (ctx) xt_ctx_destroy(ctx);
  return 0;
}
