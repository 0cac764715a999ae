// Compile-and-link check of the two drop-in headers against the reference's
// own types and call shapes: the reference's Graph (graph.hpp, included first
// as main.cpp:13-15 does), TriangleCountingSingle built from it
// (main.cpp:148-149,161,192), and the counting.cuh launchers called as the
// reference engine calls them (xxtric-gpu.hpp:195-263): the per-edge launcher
// from an OpenMP loop with the ADDRESS of one of eight streams, the masked
// launcher with and without a stream, the whole-graph launcher.
#include <unistd.h>

#include <cassert>
#include <cstdio>
#include <iostream>

#include <hip/hip_runtime_api.h>

// The reference's headers rely on main.cpp's includes above (main.cpp:1-11).
#include "deftypes.hpp"
#include "utils.hpp"
#include "graph.hpp"

#include "counting_xtric.hpp"
#include "xxtric_single.hpp"

static GraphElem engine_count(Graph* g) {
  TriangleCountingSingle tr(g);
  const GraphElem n = tr.count();
  tr.clear();
  return n;
}

// d_eind/d_elst: the graph's int64 CSR on the device; d_mask: (e1, m) pairs.
static GraphElem launcher_counts(Graph* g, const GraphElem* d_eind, const GraphElem* d_elst,
                                 const GraphElem* d_mask, GraphElem nmask, GraphElem* d_cnt) {
  constexpr int kStreams = 8;
  hipStream_t* streams = new hipStream_t[kStreams];
  for (int t = 0; t < kStreams; t++) (void)hipStreamCreateWithFlags(&streams[t], hipStreamNonBlocking);
  GraphElem nst = 0;
  const GraphElem nv = g->get_nv();
#pragma omp parallel for
  for (GraphElem i = 0; i < nv; i++) {
    GraphElem e0, e1, jstart, jend, id;
    g->edge_range(i, e0, e1);
    for (GraphElem m = e0; m + 1 < e1; m++) {
      const Edge& em = g->get_edge(m);
      if (i >= em.tail_) continue;
      g->edge_range(em.tail_, jstart, jend);
#pragma omp atomic capture
      id = nst++;
      local_count_gpu_inner(m + 1, e1, jstart, jend, d_elst, d_cnt, &streams[id % kStreams]);
    }
  }
  single_gpu_count_masked(nmask, d_mask, d_eind, d_elst, d_cnt, streams[0]);
  single_gpu_count_masked(nmask, d_mask, d_eind, d_elst, d_cnt);
  single_gpu_count(nv, d_eind, d_elst, d_cnt);
  for (int t = 0; t < kStreams; t++) (void)hipStreamSynchronize(streams[t]);
  for (int t = 0; t < kStreams; t++) (void)hipStreamDestroy(streams[t]);
  delete[] streams;
  GraphElem h = 0;
  (void)hipMemcpy(&h, d_cnt, sizeof(h), hipMemcpyDeviceToHost);
  return h;
}

int main(int argc, char* argv[]) {
  MPI_Init(&argc, &argv);
  GenerateRGG gr(1024);
  Graph* g = gr.generate(true, true, 0.0);
  const GraphElem a = engine_count(g);
  GraphElem* d = nullptr;
  const GraphElem b = (argc > 1) ? launcher_counts(g, d, d, d, 0, d) : 0;
  std::printf("%lld %lld\n", (long long)a, (long long)b);
  delete g;
  MPI_Finalize();
  return 0;
}
