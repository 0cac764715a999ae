// oracle.cpp -- CPU restatement of the xxtric counting path (TEST INFRASTRUCTURE).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, and only as the checker.  The product never links it.
//
// Compiled with -ffp-contract=off: the RGG distance test must round exactly as
// the reference's host build does (x86-64 baseline ISA, no FMA contraction).
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// ---------------------------------------------------------------------------
// Reference constants (deftypes.hpp:7,14-16).
static const double kPi = 3.14159;
static const int64_t kMlcg = 2147483647;  // 2^31 - 1
static const int64_t kAlcg = 16807;       // 7^5
static const int64_t kBlcg = 0;

extern "C" uint32_t or_reseeder(uint32_t seed) {
  // utils.hpp:37-44: std::seed_seq{seed}, one generated 32-bit word.
  std::seed_seq seq({seed});
  std::vector<std::uint32_t> out(1);
  seq.generate(out.begin(), out.end());
  return out[0];
}

extern "C" double or_rgg_radius_ref(int64_t nv) {
  // graph.hpp:789-791, same casts and operation order.
  double rc = std::sqrt((double)std::log(nv) / (double)(kPi * nv));
  double rt = std::sqrt((double)2.0736 / (double)nv);
  return (rc + rt) / (double)2.0;
}

extern "C" void or_rgg_lcg_coords(int64_t n, uint32_t seed, double* X, double* Y) {
  // utils.hpp:248-268 (x0 = reseeder(seed) at rank 0, nprocs = 1 so no
  // parallel prefix), 365-394 (generate), 399-420 (rescale with lo = 0,
  // range = 1/nprocs = 1).  The LCG object holds 2n numbers: X | Y.
  const int64_t len = 2 * n;
  std::vector<int64_t> r(len);
  r[0] = (int64_t)or_reseeder(seed);
  for (int64_t i = 1; i < len; i++) r[i] = (r[i - 1] * kAlcg + kBlcg) % kMlcg;
  const double mult = 1.0 / (double)(1.0 + (double)(kMlcg - 1));
  const double range = 1.0 / (double)1;
  for (int64_t i = 0; i < n; i++) X[i] = (double)((double)std::fabs((double)r[i]) * mult);
  for (int64_t j = 0; j < n; j++) {
    double d = (double)((double)std::fabs((double)r[n + j]) * mult);
    Y[j] = 0.0 + (double)(range * d);
  }
}

extern "C" int64_t or_rgg_lcg(int64_t n, uint32_t seed, double radius, int64_t* off,
                              int64_t* tails, int64_t cap, int nthreads) {
  return or_rgg_lcg_opts(n, seed, radius, 0, off, tails, cap, nthreads);
}

extern "C" int64_t or_rgg_lcg_opts(int64_t n, uint32_t seed, double radius, int strict, int64_t* off,
                                   int64_t* tails, int64_t cap, int nthreads) {
  // graph.hpp:919-948: every pair i<j with sqrt(dx*dx+dy*dy) <= rn_ becomes
  // the two directed entries (i,j) and (j,i); 1289-1341: CSR sorted by (src,dst).
  // strict != 0 tests < instead of <= (no counterpart in the reference).
  std::vector<double> X(n), Y(n);
  or_rgg_lcg_coords(n, seed, X.data(), Y.data());
  const double rn = radius > 0 ? radius : or_rgg_radius_ref(n);
  std::vector<std::vector<int64_t>> up(n);
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t i = 0; i < n; i++) {
    for (int64_t j = i + 1; j < n; j++) {
      double dx = X[i] - X[j];
      double dy = Y[i] - Y[j];
      double ed = std::sqrt(dx * dx + dy * dy);
      if (strict ? ed < rn : ed <= rn) up[i].push_back(j);
    }
  }
  std::vector<int64_t> deg(n, 0);
  for (int64_t i = 0; i < n; i++) {
    deg[i] += (int64_t)up[i].size();
    for (int64_t j : up[i]) deg[j]++;
  }
  off[0] = 0;
  for (int64_t i = 0; i < n; i++) off[i + 1] = off[i] + deg[i];
  const int64_t ne = off[n];
  if (tails && ne <= cap) {
    // Row j holds {i < j} in increasing i, then {j' > j} in increasing j'.
    std::vector<int64_t> cur(off, off + n);
    for (int64_t i = 0; i < n; i++)
      for (int64_t j : up[i]) tails[cur[j]++] = i;
    for (int64_t i = 0; i < n; i++)
      for (int64_t j : up[i]) tails[cur[i]++] = j;
  }
  return ne;
}

// ---------------------------------------------------------------------------
// The RGG of a P-rank run (graph.hpp:745-1374 with utils.hpp:245-430), all
// ranks' rows in one CSR.  Follows the reference step by step: each rank's
// LCG stream starts from the parallel prefix of 2x2 matrix powers computed in
// wrapping 64-bit integer arithmetic (mat_power does not reduce mod MLCG,
// utils.hpp:271-303, so the products overflow exactly as the reference's do),
// then the local pair loop and the two cross-rank loops, whose inner index
// runs only over j > i (graph.hpp:983-1066).

typedef uint64_t Mat2[4];  // row-major 2x2, wrapping arithmetic

static void mat_mul_wrap(Mat2 c, const Mat2 a, const Mat2 b) {
  // utils.hpp:271-283 matmat_2x2
  Mat2 t;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      uint64_t sum = 0;
      for (int k = 0; k < 2; k++) sum += a[i * 2 + k] * b[k * 2 + j];
      t[i * 2 + j] = sum;
    }
  std::memcpy(c, t, sizeof(Mat2));
}

static void or_rank_first_numbers(int64_t len, int nprocs, int64_t x0, std::vector<int64_t>& first) {
  // utils.hpp:248-268 (x0 broadcast) and 308-348 (parallel_prefix_op), all
  // ranks simulated together: mat_power multiplies by M (k - 1) times.
  std::vector<std::array<uint64_t, 4>> glob(nprocs), pref(nprocs);
  Mat2 m = {(uint64_t)kAlcg, 0, (uint64_t)kBlcg, 1};
  Mat2 g;
  std::memcpy(g, m, sizeof(Mat2));
  for (int64_t p = 0; p < len - 1; p++) mat_mul_wrap(g, g, m);
  for (int r = 0; r < nprocs; r++) {
    std::memcpy(glob[r].data(), g, sizeof(Mat2));
    pref[r] = {1, 0, 0, 1};
  }
  const int steps = (int)(std::log2((double)nprocs));
  for (int s = 0; s < steps; s++) {
    std::vector<std::array<uint64_t, 4>> recv(nprocs);
    for (int r = 0; r < nprocs; r++) recv[r] = glob[r ^ (1 << s)];  // MPI_Sendrecv with the mate
    for (int r = 0; r < nprocs; r++) {
      mat_mul_wrap(glob[r].data(), glob[r].data(), recv[r].data());
      if ((r ^ (1 << s)) < r) mat_mul_wrap(pref[r].data(), pref[r].data(), recv[r].data());
    }
  }
  first.assign(nprocs, 0);
  for (int r = 0; r < nprocs; r++) {
    if (r == 0) {
      first[r] = x0;
    } else {
      const int64_t t = (int64_t)((uint64_t)x0 * pref[r][0] + pref[r][2]);
      first[r] = t % kMlcg;  // C++ remainder: the sign of t
    }
  }
}

extern "C" void or_rgg_lcg_coords_ranks(int64_t n, int nprocs, uint32_t seed, double* X, double* Y) {
  // Coordinates of all ranks, rank r's points at [r*n/P, (r+1)*n/P):
  // utils.hpp:352-394 (generate), 399-420 (rescale to [r/P, (r+1)/P)),
  // graph.hpp:836-872.
  const int64_t nl = n / nprocs, len = 2 * nl;
  std::vector<int64_t> first;
  or_rank_first_numbers(len, nprocs, (int64_t)or_reseeder(seed), first);
  const double mult = 1.0 / (double)(1.0 + (double)(kMlcg - 1));
  const double rec_np = (double)(1.0 / (double)nprocs);
  const double range = (1.0 / (double)nprocs);
  std::vector<int64_t> r(len);
  for (int p = 0; p < nprocs; p++) {
    r[0] = first[p];
    for (int64_t i = 1; i < len; i++) r[i] = (r[i - 1] * kAlcg + kBlcg) % kMlcg;
    const double lo = p * rec_np;
    for (int64_t i = 0; i < nl; i++) X[p * nl + i] = (double)((double)std::fabs((double)r[i]) * mult);
    for (int64_t j = 0; j < nl; j++) {
      const double d = (double)((double)std::fabs((double)r[nl + j]) * mult);
      Y[p * nl + j] = lo + (double)(range * d);
    }
  }
}

extern "C" int64_t or_rgg_lcg_ranks(int64_t n, int nprocs, uint32_t seed, double radius, int64_t* off,
                                    int64_t* tails, int64_t cap, int nthreads) {
  // graph.hpp:919-948 (local pairs), 983-1066 (cross pairs with the rank
  // above and below: j on the neighbour runs over j > i only), 1073-1097 (the
  // received cross edges), 1289-1341 (rows sorted by neighbour id).
  if (nprocs < 1 || (nprocs & (nprocs - 1)) != 0 || n % nprocs != 0) return -1;
  const int64_t nl = n / nprocs;
  std::vector<double> X(n), Y(n);
  or_rgg_lcg_coords_ranks(n, nprocs, seed, X.data(), Y.data());
  const double rn = radius > 0 ? radius : or_rgg_radius_ref(n);
  if (!((1.0 / (double)nprocs) > rn)) return -1;  // graph.hpp:793 assert
  std::vector<std::vector<int64_t>> rows(n);
  if (nthreads < 1) nthreads = 1;
  for (int p = 0; p < nprocs; p++) {
    const int64_t b = p * nl;
    // local pairs: both entries
    std::vector<std::vector<int64_t>> up(nl);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t i = 0; i < nl; i++)
      for (int64_t j = i + 1; j < nl; j++) {
        double dx = X[b + i] - X[b + j];
        double dy = Y[b + i] - Y[b + j];
        double ed = std::sqrt(dx * dx + dy * dy);
        if (ed <= rn) up[i].push_back(j);
      }
    for (int64_t i = 0; i < nl; i++)
      for (int64_t j : up[i]) {
        rows[b + i].push_back(b + j);
        rows[b + j].push_back(b + i);
      }
    // cross pairs with the rank above (p - 1) and below (p + 1): an entry at
    // the local vertex and the mirrored entry sent to the neighbour
    for (int q : {p - 1, p + 1}) {
      if (q < 0 || q >= nprocs) continue;
      const int64_t c = q * nl;
      std::vector<std::vector<int64_t>> hit(nl);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
      for (int64_t i = 0; i < nl; i++)
        for (int64_t j = i + 1; j < nl; j++) {
          double dx = X[b + i] - X[c + j];
          double dy = Y[b + i] - Y[c + j];
          double ed = std::sqrt(dx * dx + dy * dy);
          if (ed <= rn) hit[i].push_back(j);
        }
      for (int64_t i = 0; i < nl; i++)
        for (int64_t j : hit[i]) {
          rows[b + i].push_back(c + j);
          rows[c + j].push_back(b + i);
        }
    }
  }
  off[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    std::sort(rows[i].begin(), rows[i].end());
    off[i + 1] = off[i] + (int64_t)rows[i].size();
  }
  const int64_t ne = off[n];
  if (tails && ne <= cap)
    for (int64_t i = 0; i < n; i++) std::copy(rows[i].begin(), rows[i].end(), tails + off[i]);
  return ne;
}

// ---------------------------------------------------------------------------
// Reference CPU count, xxtric.hpp:60-208.
template <typename T>
static inline bool check_edgelist(const int64_t* off, const T* tails, int64_t v, int64_t w) {
  // xxtric.hpp:516-545: first/last shortcut, then bisection with MID().
  int64_t beg = off[v], end = off[v + 1];
  int64_t mid = beg + (end - beg) / 2;
  const int64_t first = tails[beg];
  const int64_t last = tails[end - 1];
  if (w == last || w == first) return true;
  if (w > first && w < last) {
    while (end > beg) {
      const int64_t mv = tails[mid];
      if (mv == w) return true;
      if (mv < w)
        beg = mid + 1;
      else
        end = mid;
      mid = beg + (end - beg) / 2;
    }
  }
  return false;
}

template <typename T>
static uint64_t count_reference(int64_t nv, const int64_t* off, const T* tails,
                                uint64_t* per_vertex, int nthreads, int64_t vstart,
                                int64_t vstride) {
  // erange_ (xxtric.hpp:60-75): min and max neighbour of each vertex; degree-1
  // rows keep (0,0).  Degree-0 rows: the reference reads the neighbouring
  // rows' entries here (get_edge(e0-1)); that value is never consulted because
  // no edge points at an isolated vertex, so this restatement leaves (0,0).
  std::vector<int64_t> erange(2 * nv, 0);
  for (int64_t i = 0; i < nv; i++) {
    const int64_t e0 = off[i], e1 = off[i + 1];
    if (e0 + 1 == e1 || e0 == e1) continue;
    erange[2 * i] = tails[e0];
    erange[2 * i + 1] = tails[e1 - 1];
  }
  if (vstride < 1) vstride = 1;
  if (nthreads < 1) nthreads = 1;
  const int64_t nsteps = vstart < nv ? (nv - vstart + vstride - 1) / vstride : 0;
  uint64_t total = 0;
  // OpenMP schedule(static) as xxtric.hpp:84; the serial branch (149-208) is
  // the same loop with one thread.  The total is summed in 64 bits (the
  // OpenMP branch's std::accumulate(..., 0) at :147 sums in int and wraps).
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : total)
  for (int64_t k = 0; k < nsteps; k++) {
    const int64_t i = vstart + k * vstride;
    const int64_t e0 = off[i], e1 = off[i + 1];
    uint64_t t = 0;
    if (e0 + 1 != e1) {
      for (int64_t m = e0; m < e1 - 1; m++) {
        const int64_t v = tails[m];
        if (i >= v) continue;  // xxtric.hpp:165 -- triangle i<v<w counted at i
        const int64_t vmin = erange[2 * v], vmax = erange[2 * v + 1];
        for (int64_t n = m + 1; n < e1; n++) {
          const int64_t w = tails[n];
          if (!(w <= vmax)) break;     // edge_within_max, xxtric.hpp:583-588
          if (!(w >= vmin)) continue;  // edge_above_min, xxtric.hpp:576-581
          if (check_edgelist(off, tails, v, w)) t += 1;
        }
      }
    }
    if (per_vertex) per_vertex[i] = t;
    total += t;
  }
  return total;
}

extern "C" uint64_t or_count_reference(int64_t nv, const int64_t* off, const int64_t* tails,
                                       uint64_t* per_vertex, int nthreads, int64_t vstart,
                                       int64_t vstride) {
  return count_reference<int64_t>(nv, off, tails, per_vertex, nthreads, vstart, vstride);
}

extern "C" uint64_t or_count_reference32(int64_t nv, const int64_t* off, const int32_t* tails,
                                         uint64_t* per_vertex, int nthreads, int64_t vstart,
                                         int64_t vstride) {
  return count_reference<int32_t>(nv, off, tails, per_vertex, nthreads, vstart, vstride);
}

// Reference count (xxtric.hpp:160-206) restricted to the listed vertices:
// per_vertex[k] receives tcount of vertices[k].  Lets tests check per-vertex
// counts of a sample of a graph too large for a full reference run.
// Per-vertex count of listed vertices (smallest-vertex credit, xxtric.hpp:165)
// by a different route from the reference loop, for vertices it is too slow
// for (hubs: its pair loop over N(v) costs deg(v)^2 searches): mark N(v), then
// for every neighbour u > v count the marked w > u in N(u) -- each triangle
// v < u < w once, at the edge (v, u).  Rows must be ascending.  Work: the sum
// of deg(u) over u in N(v); the loop over N(v) runs on nthreads threads.
extern "C" void or_vertex_min_credit32(int64_t nv, const int64_t* off, const int32_t* tails,
                                       const int64_t* vertices, int64_t nlist, uint64_t* out,
                                       int nthreads) {
  std::vector<uint8_t> mark((size_t)nv, 0);
  if (nthreads < 1) nthreads = 1;
  for (int64_t k = 0; k < nlist; k++) {
    const int64_t v = vertices[k];
    const int64_t e0 = off[v], e1 = off[v + 1];
    for (int64_t e = e0; e < e1; e++) mark[(size_t)tails[e]] = 1;
    uint64_t c = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : c) num_threads(nthreads)
    for (int64_t e = e0; e < e1; e++) {
      const int32_t u = tails[e];
      if ((int64_t)u <= v) continue;
      const int32_t* r0 = tails + off[u];
      const int32_t* r1 = tails + off[u + 1];
      for (const int32_t* q = std::upper_bound(r0, r1, u); q < r1; q++) c += mark[(size_t)*q];
    }
    for (int64_t e = e0; e < e1; e++) mark[(size_t)tails[e]] = 0;
    out[k] = c;
  }
}

extern "C" void or_count_reference32_list(int64_t nv, const int64_t* off, const int32_t* tails,
                                          const int64_t* vertices, int64_t nlist,
                                          uint64_t* per_vertex, int nthreads) {
  // erange of the neighbours the listed rows reach, as xxtric.hpp:60-75
  auto vmin = [&](int64_t v) -> int64_t {
    const int64_t e0 = off[v], e1 = off[v + 1];
    return (e0 + 1 == e1 || e0 == e1) ? 0 : tails[e0];
  };
  auto vmax = [&](int64_t v) -> int64_t {
    const int64_t e0 = off[v], e1 = off[v + 1];
    return (e0 + 1 == e1 || e0 == e1) ? 0 : tails[e1 - 1];
  };
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t k = 0; k < nlist; k++) {
    const int64_t i = vertices[k];
    const int64_t e0 = off[i], e1 = off[i + 1];
    uint64_t t = 0;
    if (e0 + 1 != e1) {
      for (int64_t m = e0; m < e1 - 1; m++) {
        const int64_t v = tails[m];
        if (i >= v) continue;
        const int64_t lo = vmin(v), hi = vmax(v);
        for (int64_t n = m + 1; n < e1; n++) {
          const int64_t w = tails[n];
          if (!(w <= hi)) break;
          if (!(w >= lo)) continue;
          if (check_edgelist(off, tails, v, w)) t += 1;
        }
      }
    }
    per_vertex[k] = t;
  }
}

// Time-bounded sample of the reference count for CPU baselines: vertices are
// taken in id order (a random order for relabelled R-MAT graphs) by nthreads
// threads until `seconds` elapse; a vertex still running at the deadline is
// abandoned and not counted.  Returns the number of vertices completed and
// their triangles; *edges_done receives the sum of their degrees.
extern "C" int64_t or_count_reference32_budget(int64_t nv, const int64_t* off, const int32_t* tails,
                                               int nthreads, double seconds, uint64_t* triangles,
                                               int64_t* edges_done) {
  std::vector<int64_t> erange(2 * nv, 0);
  for (int64_t i = 0; i < nv; i++) {
    const int64_t e0 = off[i], e1 = off[i + 1];
    if (e0 + 1 == e1 || e0 == e1) continue;
    erange[2 * i] = tails[e0];
    erange[2 * i + 1] = tails[e1 - 1];
  }
  if (nthreads < 1) nthreads = 1;
  const double t0 = omp_get_wtime();
  int64_t next = 0, done = 0, edges = 0;
  uint64_t tri = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : done, edges, tri)
  {
    bool stop = false;
    while (!stop) {
      int64_t i;
#pragma omp atomic capture
      i = next++;
      if (i >= nv) break;
      const int64_t e0 = off[i], e1 = off[i + 1];
      uint64_t t = 0;
      int64_t steps = 0;
      bool aborted = false;
      if (e0 + 1 != e1) {
        for (int64_t m = e0; m < e1 - 1 && !aborted; m++) {
          const int64_t v = tails[m];
          if (i >= v) continue;
          const int64_t vmin = erange[2 * v], vmax = erange[2 * v + 1];
          for (int64_t n = m + 1; n < e1; n++) {
            const int64_t w = tails[n];
            if (!(w <= vmax)) break;
            if (!(w >= vmin)) continue;
            if (check_edgelist(off, tails, v, w)) t += 1;
            if ((++steps & 4095) == 0 && omp_get_wtime() - t0 > seconds) {
              aborted = true;
              break;
            }
          }
        }
      }
      if (aborted) break;
      tri += t;
      done += 1;
      edges += e1 - e0;
      if ((done & 255) == 0 && omp_get_wtime() - t0 > seconds) stop = true;
    }
  }
  *triangles = tri;
  *edges_done = edges;
  return done;
}

// ---------------------------------------------------------------------------
// Independent exact counter (not a reference restatement): orientation by
// (degree, id), sorted-merge intersection of out-lists, credit min(u, v, w).
extern "C" uint64_t or_count_degree_ordered(int64_t nv, const int64_t* off, const int32_t* tails,
                                            uint64_t* per_vertex, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  auto deg = [&](int64_t x) { return off[x + 1] - off[x]; };
  auto less = [&](int64_t a, int64_t b) {
    const int64_t da = deg(a), db = deg(b);
    return da < db || (da == db && a < b);
  };
  std::vector<int64_t> ooff(nv + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
  for (int64_t u = 0; u < nv; u++) {
    int64_t c = 0;
    for (int64_t e = off[u]; e < off[u + 1]; e++)
      if (less(u, tails[e])) c++;
    ooff[u + 1] = c;
  }
  for (int64_t u = 0; u < nv; u++) ooff[u + 1] += ooff[u];
  std::vector<int32_t> ot(ooff[nv]);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
  for (int64_t u = 0; u < nv; u++) {
    int64_t p = ooff[u];
    for (int64_t e = off[u]; e < off[u + 1]; e++)
      if (less(u, tails[e])) ot[p++] = tails[e];
  }
  if (per_vertex) std::memset(per_vertex, 0, sizeof(uint64_t) * nv);
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) reduction(+ : total)
  for (int64_t u = 0; u < nv; u++) {
    for (int64_t a = ooff[u]; a < ooff[u + 1]; a++) {
      const int64_t v = ot[a];
      int64_t i = ooff[u], ie = ooff[u + 1], j = ooff[v], je = ooff[v + 1];
      const int64_t uv = u < v ? u : v;
      uint64_t at_uv = 0;
      while (i < ie && j < je) {
        const int64_t x = ot[i], y = ot[j];
        if (x < y)
          i++;
        else if (y < x)
          j++;
        else {
          total++;
          if (per_vertex) {
            if (x < uv) {
#pragma omp atomic
              per_vertex[x] += 1;
            } else {
              at_uv++;
            }
          }
          i++;
          j++;
        }
      }
      if (per_vertex && at_uv) {
#pragma omp atomic
        per_vertex[uv] += at_uv;
      }
    }
  }
  return total;
}

// ---------------------------------------------------------------------------
// vite binary format, graph.hpp:294-395 (reader) with the layout at 321-323,
// 333-334, 361-362; Edge = {GraphElem tail_; GraphWeight weight_} (graph.hpp:20-25).
template <typename E, typename W>
static int vite_write(const char* path, int64_t nv, const int64_t* off, const int64_t* tails) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return -1;
  E hdr[2] = {(E)nv, (E)off[nv]};
  int ok = std::fwrite(hdr, sizeof(E), 2, f) == 2;
  std::vector<E> o(nv + 1);
  for (int64_t i = 0; i <= nv; i++) o[i] = (E)off[i];
  ok = ok && std::fwrite(o.data(), sizeof(E), nv + 1, f) == (size_t)(nv + 1);
  struct Rec {
    E tail;
    W w;
  };
  const int64_t ne = off[nv];
  std::vector<Rec> buf(1 << 16);
  for (int64_t e = 0; e < ne && ok;) {
    int64_t k = std::min<int64_t>(ne - e, (int64_t)buf.size());
    for (int64_t t = 0; t < k; t++) {
      buf[t].tail = (E)tails[e + t];
      buf[t].w = (W)1.0;
    }
    ok = std::fwrite(buf.data(), sizeof(Rec), k, f) == (size_t)k;
    e += k;
  }
  std::fclose(f);
  return ok ? 0 : -1;
}

template <typename E, typename W>
static int vite_read(const char* path, int64_t* nv, int64_t* ne, int64_t* off, int64_t* tails) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return -1;
  E hdr[2];
  if (std::fread(hdr, sizeof(E), 2, f) != 2) {
    std::fclose(f);
    return -2;
  }
  *nv = hdr[0];
  *ne = hdr[1];
  if (!off) {
    std::fclose(f);
    return 0;
  }
  std::vector<E> o(*nv + 1);
  if (std::fread(o.data(), sizeof(E), *nv + 1, f) != (size_t)(*nv + 1)) {
    std::fclose(f);
    return -3;
  }
  for (int64_t i = 0; i <= *nv; i++) off[i] = (int64_t)o[i] - (int64_t)o[0];
  struct Rec {
    E tail;
    W w;
  };
  std::vector<Rec> buf(1 << 16);
  for (int64_t e = 0; e < *ne;) {
    int64_t k = std::min<int64_t>(*ne - e, (int64_t)buf.size());
    if (std::fread(buf.data(), sizeof(Rec), k, f) != (size_t)k) {
      std::fclose(f);
      return -4;
    }
    for (int64_t t = 0; t < k; t++) tails[e + t] = (int64_t)buf[t].tail;
    e += k;
  }
  std::fclose(f);
  return 0;
}

extern "C" int or_vite_write(const char* path, int64_t nv, const int64_t* off,
                             const int64_t* tails, int layout32) {
  return layout32 ? vite_write<int32_t, float>(path, nv, off, tails)
                  : vite_write<int64_t, double>(path, nv, off, tails);
}

extern "C" int or_vite_read(const char* path, int layout32, int64_t* nv, int64_t* ne,
                            int64_t* off, int64_t* tails) {
  return layout32 ? vite_read<int32_t, float>(path, nv, ne, off, tails)
                  : vite_read<int64_t, double>(path, nv, ne, off, tails);
}

// ---------------------------------------------------------------------------
// This build's own generators.  The mixing function is the SplitMix64 output
// function (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014; constants as in Vigna's public splitmix64.c).
extern "C" uint64_t or_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

extern "C" void or_rmat_thresholds(double a, double b, double c, uint32_t* t) {
  const double s = 4294967296.0;
  t[0] = (uint32_t)std::floor(a * s);
  t[1] = (uint32_t)std::floor((a + b) * s);
  t[2] = (uint32_t)std::floor((a + b + c) * s);
}

static inline uint32_t rmat_perm(uint32_t x, int scale, uint64_t key) {
  // Bijection on [0, 2^scale): three rounds of odd multiply + add mod 2^scale
  // followed by an xor-shift right (each step is invertible on scale bits).
  const uint64_t mask = (scale >= 64) ? ~0ULL : ((1ULL << scale) - 1);
  const int h = (scale + 1) / 2;
  uint64_t y = x;
  for (int r = 0; r < 3; r++) {
    const uint64_t m = or_mix64(key + 1000 + r) | 1ULL;
    const uint64_t k = or_mix64(key + 2000 + r);
    y = (y * m + k) & mask;
    y ^= y >> h;
  }
  return (uint32_t)y;
}

extern "C" void or_rmat_tuples(int scale, uint64_t seed, const uint32_t* thr, int permute,
                               int64_t start, int64_t count, uint32_t* src, uint32_t* dst) {
  const uint64_t key = or_mix64(seed);
  for (int64_t t = 0; t < count; t++) {
    const uint64_t e = (uint64_t)(start + t);
    uint32_t u = 0, v = 0;
    uint64_t w = 0;
    for (int l = 0; l < scale; l++) {
      if ((l & 1) == 0) w = or_mix64(key + e * 32 + (uint64_t)(l >> 1));
      const uint32_t r = (l & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
      uint32_t ub, vb;
      if (r < thr[0]) {
        ub = 0; vb = 0;
      } else if (r < thr[1]) {
        ub = 0; vb = 1;
      } else if (r < thr[2]) {
        ub = 1; vb = 0;
      } else {
        ub = 1; vb = 1;
      }
      u = (u << 1) | ub;
      v = (v << 1) | vb;
    }
    if (permute) {
      u = rmat_perm(u, scale, key);
      v = rmat_perm(v, scale, key);
    }
    src[t] = u;
    dst[t] = v;
  }
}

extern "C" int64_t or_rmat_csr(int scale, int edgefactor, uint64_t seed, const uint32_t* thr,
                               int permute, int64_t* off, int32_t* tails, int64_t cap) {
  const int64_t nv = 1LL << scale;
  const int64_t nt = (int64_t)edgefactor * nv;
  std::vector<uint32_t> s(nt), d(nt);
  or_rmat_tuples(scale, seed, thr, permute, 0, nt, s.data(), d.data());
  std::vector<uint64_t> keys;
  keys.reserve(2 * nt);
  for (int64_t t = 0; t < nt; t++) {
    if (s[t] == d[t]) continue;
    keys.push_back(((uint64_t)s[t] << 32) | d[t]);
    keys.push_back(((uint64_t)d[t] << 32) | s[t]);
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  std::fill(off, off + nv + 1, 0);
  for (uint64_t k : keys) off[(k >> 32) + 1]++;
  for (int64_t i = 0; i < nv; i++) off[i + 1] += off[i];
  const int64_t ne = (int64_t)keys.size();
  if (tails && ne <= cap)
    for (int64_t e = 0; e < ne; e++) tails[e] = (int32_t)(keys[e] & 0xffffffffULL);
  return ne;
}

extern "C" int64_t or_er_csr(int64_t n, uint64_t seed, uint64_t thr64, int64_t* off,
                             int32_t* tails, int64_t cap) {
  const uint64_t key = or_mix64(seed);
  std::vector<std::vector<int32_t>> up(n);
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = i + 1; j < n; j++)
      if ((or_mix64(key + (uint64_t)i * (uint64_t)n + (uint64_t)j) >> 32) < thr64)
        up[i].push_back((int32_t)j);
  std::vector<int64_t> deg(n, 0);
  for (int64_t i = 0; i < n; i++) {
    deg[i] += (int64_t)up[i].size();
    for (int32_t j : up[i]) deg[j]++;
  }
  off[0] = 0;
  for (int64_t i = 0; i < n; i++) off[i + 1] = off[i] + deg[i];
  const int64_t ne = off[n];
  if (tails && ne <= cap) {
    std::vector<int64_t> cur(off, off + n);
    for (int64_t i = 0; i < n; i++)
      for (int32_t j : up[i]) tails[cur[j]++] = (int32_t)i;
    for (int64_t i = 0; i < n; i++)
      for (int32_t j : up[i]) tails[cur[i]++] = j;
  }
  return ne;
}

extern "C" void or_batch_params(uint64_t base_seed, int64_t g, int64_t nmin, int64_t nmax,
                                int64_t* n, int* kind, uint64_t* gseed, uint64_t* er_thr64) {
  const uint64_t gk = or_mix64(or_mix64(base_seed) + (uint64_t)g);
  const double u1 = (double)(or_mix64(gk + 1) >> 11) * 0x1.0p-53;
  const double ln0 = std::log((double)nmin), ln1 = std::log((double)nmax);
  int64_t nn = (int64_t)std::llround(std::exp(ln0 + u1 * (ln1 - ln0)));
  if (nn < nmin) nn = nmin;
  if (nn > nmax) nn = nmax;
  *n = nn;
  *kind = (int)(g & 1);
  *gseed = gk;
  const double u2 = (double)(or_mix64(gk + 2) >> 11) * 0x1.0p-53;
  const double deg = 4.0 + 60.0 * u2;
  const double p = nn > 1 ? deg / (double)(nn - 1) : 0.0;
  *er_thr64 = p >= 1.0 ? (1ULL << 32) : (uint64_t)std::floor(p * 4294967296.0);
}

// ---- vertex partitions (BinaryEdgeList::read_balanced, graph.hpp:608-640) ----

// graph.hpp:556-602 (find_balanced_num_edges, the default build's -b): walk the
// vertices in order; a vertex joins the current rank while the rank stays at
// or under ne / nparts entries (the last rank takes everything left), else it
// opens the next rank.  bins are counts per rank, then a prefix sum.
extern "C" int or_partition_greedy(int64_t nv, const int64_t* off, int64_t nparts,
                                   int64_t* bins) {
  if (nv < 0 || nparts < 1) return -1;
  for (int64_t r = 0; r <= nparts; r++) bins[r] = 0;
  std::vector<int64_t> held(nparts, 0);
  const int64_t cap = (off[nv] - off[0]) / nparts;
  int64_t r = 0;
  for (int64_t v = 0; v < nv; v++) {
    const int64_t d = off[v + 1] - off[v];
    if (held[r] + d > cap && r != nparts - 1) r++;
    held[r] += d;
    bins[r + 1]++;
  }
  for (int64_t q = 1; q <= nparts; q++) bins[q] += bins[q - 1];
  return 0;
}

// graph.hpp:397-419 + utils.hpp:46-135 (-d).  The sequence is the first nv
// offsets.  best[i][j]: smallest possible largest part when seq[0..i] is cut
// into j+1 consecutive parts, with best[i][0] the prefix sum and best[0][j] =
// seq[0]; cut[i][j] is the first split x (the last part is seq[x+1..i]) that
// attains best[i][j], scanning x = 0..i-1 and keeping only strict improvements.
// The split points are then walked back from the last part.
extern "C" int or_partition_linear(int64_t nv, const int64_t* off, int64_t nparts,
                                   int64_t* bins) {
  const int64_t n = nv, k = nparts;
  if (n < 1 || k < 1) return -1;
  if (k > n - 1) return -1;  // the reference fills bins with 1s and goes on: not a partition
  for (int64_t r = 0; r <= k; r++) bins[r] = 0;
  std::vector<int64_t> best((size_t)(n * k)), cut((size_t)((n - 1) * (k - 1)));
  auto B = [&](int64_t i, int64_t j) -> int64_t& { return best[(size_t)(i * k + j)]; };
  int64_t run = 0;
  for (int64_t i = 0; i < n; i++) {
    run += off[i];
    B(i, 0) = run;
  }
  for (int64_t j = 0; j < k; j++) B(0, j) = off[0];
  for (int64_t i = 1; i < n; i++) {
    for (int64_t j = 1; j < k; j++) {
      int64_t lo = 0, at = 0;
      for (int64_t x = 0; x < i; x++) {
        const int64_t last = B(i, 0) - B(x, 0);
        const int64_t worst = B(x, j - 1) > last ? B(x, j - 1) : last;
        if (x == 0 || worst < lo) {
          lo = worst;
          at = x;
        }
      }
      B(i, j) = lo;
      cut[(size_t)((i - 1) * (k - 1) + (j - 1))] = at;
    }
  }
  // Walk back: rank p takes seq[cut+1 .. i]; the first rank takes seq[0 .. i].
  int64_t i = n - 1, p = k;
  for (int64_t j = k - 2; j >= 0; j--, p--) {
    const int64_t idx = (i - 1) * (k - 1) + j;
    if (i < 1 || idx < 0 || idx >= (n - 1) * (k - 1)) return -1;  // outside the reference's table
    const int64_t x = cut[(size_t)idx];
    bins[p] += i - x;
    i = x;
  }
  bins[p] += i + 1;
  for (int64_t q = 1; q <= k; q++) bins[q] += bins[q - 1];
  return 0;
}
