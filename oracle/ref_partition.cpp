// ref_partition.cpp -- TEST INFRASTRUCTURE ONLY: prints the vertex bins that
// the reference's own routines compute for a vite file, so the oracle's
// restatement can be checked against them.  Built by `make -C oracle ref`
// against the reference's headers where they lie (graph.hpp, utils.hpp,
// deftypes.hpp); nothing of them is copied here.
//
//   ref_partition FILE NPARTS METHOD     METHOD 0: BinaryEdgeList::find_balanced_num_edges
//                                        (graph.hpp:556-602, flag -b);
//                                        METHOD 1: linear_partitioning_num_edges
//                                        (graph.hpp:397-419, flag -d)
// Output: one line, the NPARTS + 1 bins separated by spaces.
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
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s FILE NPARTS METHOD\n", argv[0]);
    return 2;
  }
  MPI_Init(&argc, &argv);
  const std::string file = argv[1];
  const int nparts = std::atoi(argv[2]);
  const int method = std::atoi(argv[3]);
  std::vector<GraphElem> bins(nparts + 1, 0);
  BinaryEdgeList reader;
  if (method == 0)
    reader.find_balanced_num_edges(nparts, file, bins);
  else
    reader.linear_partitioning_num_edges(nparts, file, bins);
  for (int r = 0; r <= nparts; r++) std::printf(r ? " %lld" : "%lld", (long long)bins[r]);
  std::printf("\n");
  MPI_Finalize();
  return 0;
}
