// sanitize_host.cpp -- drives the library's host-only code (ingest.cpp,
// partition.cpp) under AddressSanitizer and UndefinedBehaviorSanitizer.
// Built and run by tests/test_sanitizers.py with g++ -fsanitize=address,undefined;
// no device code is involved.  Exit status 0 when every call returned what it
// should; a sanitizer report aborts with a nonzero status.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "xtric.h"

#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      return 1;                                                      \
    }                                                                \
  } while (0)

static int rgg(int64_t n, int p, std::vector<int64_t>& off, std::vector<int64_t>& tails) {
  int64_t ne = 0;
  off.assign(n + 1, 0);
  if (xt_rgg_generate_ranks(n, p, 1, 0.0, off.data(), nullptr, 0, &ne) != XT_OK) return 1;
  tails.assign(ne > 0 ? ne : 1, 0);
  return xt_rgg_generate_ranks(n, p, 1, 0.0, off.data(), tails.data(), ne, &ne) != XT_OK;
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  std::vector<int64_t> off, tails;
  // generators: one rank and several, and the cases the reference aborts on
  CHECK(rgg(4096, 1, off, tails) == 0 && off[4096] == 29344);
  CHECK(rgg(4096, 4, off, tails) == 0 && off[4096] == 29438);
  CHECK(rgg(3000, 8, off, tails) == 0);
  CHECK(xt_rgg_generate_ranks(4096, 3, 1, 0.0, off.data(), nullptr, 0, nullptr) == XT_ERR_ARG);
  int64_t ne = 0;
  CHECK(xt_rgg_generate_ranks(4096, 256, 1, 0.0, off.data(), nullptr, 0, &ne) == XT_ERR_ARG);
  CHECK(xt_rgg_generate(1, 1, 0.0, off.data(), nullptr, 0, &ne) == XT_OK && ne == 0);
  // vite files: both layouts round trip; truncated files are rejected
  CHECK(rgg(2048, 2, off, tails) == 0);
  const int64_t nv = 2048, m = off[nv];
  for (int l32 = 0; l32 < 2; l32++) {
    const std::string path = dir + "/g" + std::to_string(l32) + ".bin";
    CHECK(xt_vite_write(path.c_str(), l32, nv, off.data(), tails.data()) == XT_OK);
    int64_t rn = 0, rm = 0;
    CHECK(xt_vite_read(path.c_str(), l32, &rn, &rm, nullptr, nullptr) == XT_OK && rn == nv && rm == m);
    std::vector<int64_t> o2(rn + 1), t2(rm);
    CHECK(xt_vite_read(path.c_str(), l32, &rn, &rm, o2.data(), t2.data()) == XT_OK);
    CHECK(o2 == off && std::vector<int64_t>(t2.begin(), t2.end()) ==
                           std::vector<int64_t>(tails.begin(), tails.begin() + m));
    std::FILE* f = std::fopen(path.c_str(), "r+b");
    CHECK(f != nullptr);
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fclose(f);
    CHECK(truncate(path.c_str(), size - 5) == 0);
    CHECK(xt_vite_read(path.c_str(), l32, &rn, &rm, o2.data(), t2.data()) == XT_ERR_INPUT);
  }
  CHECK(xt_vite_read((dir + "/missing.bin").c_str(), 0, &ne, &ne, nullptr, nullptr) != XT_OK);
  // edge-list text files
  {
    const std::string path = dir + "/edges.tsv";
    std::FILE* f = std::fopen(path.c_str(), "w");
    CHECK(f != nullptr);
    std::fprintf(f, "1\t2\t1.0\n2\t3\t1.0\n3\t1\t1.0\n3\t1\t1.0\n4\t4\t1.0\n");
    std::fclose(f);
    int64_t en = 0, em = 0;
    CHECK(xt_edgelist_to_csr(path.c_str(), &en, &em, nullptr, nullptr) == XT_OK);
    std::vector<int64_t> eo(en + 1), et(em > 0 ? em : 1);
    CHECK(xt_edgelist_to_csr(path.c_str(), &en, &em, eo.data(), et.data()) == XT_OK);
  }
  // vertex partitions: both methods, several rank counts, bad inputs
  std::vector<int64_t> bins(65);
  for (int parts : {1, 2, 3, 8, 64}) {
    CHECK(xt_partition_vertices(nv, off.data(), parts, XT_PARTITION_GREEDY, bins.data()) == XT_OK);
    CHECK(bins[0] == 0 && bins[parts] == nv);
    if (parts <= 8) CHECK(xt_partition_vertices(nv, off.data(), parts, XT_PARTITION_LINEAR, bins.data()) == XT_OK);
  }
  std::vector<int64_t> down = {0, 5, 3, 9};
  CHECK(xt_partition_vertices(3, down.data(), 2, XT_PARTITION_GREEDY, bins.data()) == XT_ERR_INPUT);
  CHECK(xt_partition_vertices(3, off.data(), 5, XT_PARTITION_LINEAR, bins.data()) == XT_ERR_ARG);
  // the batch mix's host generator
  for (int64_t g = 0; g < 6; g++) {
    int64_t bn = 0, bm = 0;
    CHECK(xt_batch_graph_csr(2026, g, 200, 2000, &bn, &bm, nullptr, nullptr) == XT_OK);
    std::vector<int64_t> bo(bn + 1);
    std::vector<int32_t> bt(bm > 0 ? bm : 1);
    CHECK(xt_batch_graph_csr(2026, g, 200, 2000, &bn, &bm, bo.data(), bt.data()) == XT_OK);
  }
  std::printf("sanitized host checks passed\n");
  return 0;
}
