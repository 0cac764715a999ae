// partition.cpp -- vertex partitions for a split of a graph over ranks (host
// code of libxtric; no device work).
//
// Before a balanced read the reference decides how many consecutive vertices
// each of nparts ranks owns (BinaryEdgeList::read_balanced, graph.hpp:608-640):
//
//  * greedy (flag -b; graph.hpp:556-602, the default build's
//    find_balanced_num_edges): vertices are taken in order; a vertex joins the
//    current rank while that rank stays at or under ne / nparts entries, else
//    it opens the next rank; the last rank takes whatever is left.
//
//  * linear partitioning (flag -d; graph.hpp:397-419 with utils.hpp:46-135):
//    Skiena's dynamic program, which minimises the largest part's sum, run on
//    seq[i] = offsets[i] for i < nv -- the first nv offsets of the file, i.e.
//    running entry counts rather than per-vertex counts, exactly as the
//    reference reads them.
//
// The reference fills an nv x nparts table in O(nv^2 nparts) time.  Here the
// same table and the same split points come out of O(nv nparts) work, one
// column at a time.  For a nonnegative sequence with prefix sums S:
//   best[i][j] = min over x < i of max(best[x][j-1], S[i] - S[x]),
// where best[x][j-1] is nondecreasing in x and S[i] - S[x] nonincreasing, so
// the first x with best[x][j-1] + S[x] >= S[i] (call it c; c = i if none)
// splits the range: below c the maximum is S[i] - S[x], smallest at c-1 and
// first reached at the first x with S[x] == S[c-1]; from c on it is
// best[x][j-1], smallest (first) at c.  The reference keeps the first x of
// the smallest value, so the earlier candidate wins ties.  c never decreases
// as i grows (S[i] does not), so one pointer per column finds every c.
#include <cstdint>
#include <vector>

#include "../../include/xtric.h"
#include "xt_guard.hpp"

namespace {

int partition_greedy(int64_t nv, const int64_t* off, int64_t k, int64_t* bins) {
  for (int64_t r = 0; r <= k; r++) bins[r] = 0;
  const int64_t cap = (off[nv] - off[0]) / k;  // the vite header's ne over the ranks
  int64_t r = 0, held = 0;
  for (int64_t v = 0; v < nv; v++) {
    const int64_t d = off[v + 1] - off[v];
    if (held + d > cap && r != k - 1) {
      r++;
      held = 0;
    }
    held += d;
    bins[r + 1]++;
  }
  for (int64_t q = 1; q <= k; q++) bins[q] += bins[q - 1];
  return XT_OK;
}

int partition_linear(int64_t n, const int64_t* seq, int64_t k, int64_t* bins) {
  // Prefix sums; a sum past 2^63 overflows the reference's GraphElem too.
  std::vector<int64_t> S((size_t)n);
  int64_t run = 0;
  for (int64_t i = 0; i < n; i++) {
    if (__builtin_add_overflow(run, seq[i], &run)) return XT_ERR_ARG;
    S[(size_t)i] = run;
  }
  // first[x]: first index whose prefix sum equals S[x].
  std::vector<int32_t> first((size_t)n);
  for (int64_t x = 0; x < n; x++)
    first[(size_t)x] = (x > 0 && S[(size_t)x] == S[(size_t)x - 1]) ? first[(size_t)x - 1] : (int32_t)x;
  // Split points, column-major: cut[(j-1)*(n-1) + (i-1)] for i >= 1, j >= 1.
  std::vector<int32_t> cut((size_t)((n - 1) * (k - 1)));
  std::vector<int64_t> prev(S), cur((size_t)n);  // column 0 is the prefix sums
  prev[0] = seq[0];
  for (int64_t j = 1; j < k; j++) {
    cur[0] = seq[0];
    int64_t c = 0;
    for (int64_t i = 1; i < n; i++) {
      const int64_t Si = S[(size_t)i];
      // best[x][j-1] + S[x] is nondecreasing: the pointer only moves forward.
      while (c < i && (__int128)prev[(size_t)c] + S[(size_t)c] < Si) c++;  // no overflow near 2^63
      int64_t val, at;
      if (c == 0) {
        val = prev[0];
        at = 0;
      } else {
        const int64_t below = Si - S[(size_t)c - 1];
        if (c == i || below <= prev[(size_t)c]) {
          val = below;
          at = first[(size_t)c - 1];
        } else {
          val = prev[(size_t)c];
          at = c;
        }
      }
      cur[(size_t)i] = val;
      cut[(size_t)((j - 1) * (n - 1) + (i - 1))] = (int32_t)at;
    }
    prev.swap(cur);
  }
  // Walk back from the last rank: rank p takes seq[x+1 .. i], the first rank seq[0 .. i].
  for (int64_t r = 0; r <= k; r++) bins[r] = 0;
  int64_t i = n - 1, p = k;
  for (int64_t j = k - 2; j >= 0; j--, p--) {
    if (i < 1) return XT_ERR_ARG;  // the reference reads outside its table here
    const int64_t x = cut[(size_t)(j * (n - 1) + (i - 1))];
    bins[p] += i - x;
    i = x;
  }
  bins[p] += i + 1;
  for (int64_t q = 1; q <= k; q++) bins[q] += bins[q - 1];
  return XT_OK;
}

}  // namespace

extern "C" {

int xt_partition_vertices(int64_t nv, const int64_t* offsets, int64_t nparts, int method,
                          int64_t* bins) {
  return xt_guarded([&]() -> int {
    if (nv < 0 || !offsets || nparts < 1 || !bins) return XT_ERR_ARG;
    if (nv >= (int64_t)INT32_MAX) return XT_ERR_ARG;
    if (offsets[0] < 0) return XT_ERR_INPUT;
    for (int64_t v = 0; v < nv; v++)
      if (offsets[v + 1] < offsets[v]) return XT_ERR_INPUT;
    if (method == XT_PARTITION_GREEDY) return partition_greedy(nv, offsets, nparts, bins);
    if (method != XT_PARTITION_LINEAR) return XT_ERR_ARG;
    // nparts > nv - 1: the reference's bins are not a partition (it fills them
    // with ones and then adds the table walk on top).
    if (nv < 1 || nparts > nv - 1) return XT_ERR_ARG;
    // Its tables take 4 (nv - 1)(nparts - 1) + 28 nv bytes of host memory
    // (XT_ERR_NOMEM from the guard when they cannot be allocated).
    return partition_linear(nv, offsets, nparts, bins);
  });
}

}  // extern "C"
