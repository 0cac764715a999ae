// ref_rgg_ranks.cpp -- TEST INFRASTRUCTURE ONLY: writes the random geometric
// graph that the reference's own generator builds when run on P MPI ranks
// (GenerateRGG::generate in LCG mode, graph.hpp:745-1374, with the LCG parallel
// prefix and per-rank rescale of utils.hpp:245-430), gathered into one CSR,
// so the oracle's restatement and the library's generator can be checked
// against it.  Built by `make -C oracle ref` against the reference's headers
// where they lie; nothing of them is copied here.  The reference's own driver
// segfaults in its distributed count at P > 1 in this container, which is why
// the graph is taken from the generator directly.
//
//   mpiexec -n P ref_rgg_ranks NV OUT
// Output file (rank 0): int64 nv, int64 ne, int64 offsets[nv + 1], int64 tails[ne]
// (rank r's rows are vertices r*nv/P .. (r+1)*nv/P - 1, ids global).
#include <mpi.h>
#include <sys/resource.h>
#include <sys/time.h>
#include <unistd.h>

#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "deftypes.hpp"
#include "utils.hpp"
#include "graph.hpp"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: mpiexec -n P %s NV OUT\n", argv[0]);
    return 2;
  }
  MPI_Init(&argc, &argv);
  int rank = 0, size = 1;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &size);
  const GraphElem nv = std::atol(argv[1]);
  GenerateRGG gr(nv);
  Graph* g = gr.generate(true, true, 0.0);
  const GraphElem lnv = g->get_lnv(), lne = g->get_lne();
  std::vector<long long> deg(lnv), tails(lne);
  for (GraphElem i = 0; i < lnv; i++) deg[i] = g->edge_indices_[i + 1] - g->edge_indices_[i];
  for (GraphElem e = 0; e < lne; e++) tails[e] = g->get_edge(e).tail_;
  // Gather the per-rank row lengths and ids on rank 0, in rank order.
  int nloc = (int)lnv, eloc = (int)lne;
  std::vector<int> nvs(size), nes(size), dv(size), de(size);
  MPI_Gather(&nloc, 1, MPI_INT, nvs.data(), 1, MPI_INT, 0, MPI_COMM_WORLD);
  MPI_Gather(&eloc, 1, MPI_INT, nes.data(), 1, MPI_INT, 0, MPI_COMM_WORLD);
  long long tv = 0, te = 0;
  if (rank == 0)
    for (int r = 0; r < size; r++) {
      dv[r] = (int)tv;
      de[r] = (int)te;
      tv += nvs[r];
      te += nes[r];
    }
  std::vector<long long> all_deg(rank == 0 ? tv : 0), all_tails(rank == 0 ? te : 0);
  MPI_Gatherv(deg.data(), nloc, MPI_LONG_LONG, all_deg.data(), nvs.data(), dv.data(), MPI_LONG_LONG, 0,
              MPI_COMM_WORLD);
  MPI_Gatherv(tails.data(), eloc, MPI_LONG_LONG, all_tails.data(), nes.data(), de.data(), MPI_LONG_LONG, 0,
              MPI_COMM_WORLD);
  if (rank == 0) {
    std::FILE* f = std::fopen(argv[2], "wb");
    if (!f) {
      std::fprintf(stderr, "cannot open %s\n", argv[2]);
      MPI_Abort(MPI_COMM_WORLD, 1);
    }
    long long hdr[2] = {tv, te};
    std::fwrite(hdr, sizeof(long long), 2, f);
    long long o = 0;
    std::fwrite(&o, sizeof(long long), 1, f);
    for (long long i = 0; i < tv; i++) {
      o += all_deg[i];
      std::fwrite(&o, sizeof(long long), 1, f);
    }
    std::fwrite(all_tails.data(), sizeof(long long), (size_t)te, f);
    std::fclose(f);
  }
  MPI_Finalize();
  return 0;
}
