/*
 * oracle.h -- CPU restatement of the xxtric counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (xxtric_amd/) links,
 * loads or calls this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do, and there only as the checker.
 *
 * Every function restates a piece of the reference (sg0/xxtric) and cites the
 * file:line it follows, or is one of this build's own generators (R-MAT, ER),
 * restated here on the host so the device generators can be checked against it.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   or_count_reference      golden counts from the reference binary (oracle/_ref),
 *                           brute force O(n^3), closed forms (K_n, bipartite, grid),
 *                           scipy row sums of (U*U).*U for per-vertex counts.
 *   or_count_degree_ordered equality with or_count_reference (global + per-vertex)
 *                           and with the closed forms above.
 *   or_vertex_min_credit32  equality with or_count_reference's per-vertex counts on
 *                           every vertex of R-MAT, RGG and ER fixtures, and the K_n
 *                           closed form C(n-1-i, 2).
 *   or_rgg_lcg              golden edge and triangle counts from the reference binary.
 *   or_rgg_lcg_opts         strict test: a radius equal to one pair's distance drops
 *                           that pair; both tests equal an all-pairs numpy enumeration.
 *   or_rgg_lcg_ranks        the CSR (sha256) the reference's generator builds on P
 *                           MPI ranks (oracle/_ref/ref_rgg_ranks under mpiexec); at
 *                           P = 1 equal to or_rgg_lcg.
 *   or_vite_read/write      byte-layout round trip and reading by the reference binary.
 *   or_partition_greedy /   bins of the reference's own routines (graph.hpp:397-602,
 *   or_partition_linear     utils.hpp:46-135) compiled from its headers (oracle/_ref/
 *                           ref_partition) on fixture offsets; the dynamic program's
 *                           largest part equals a brute-force minimum over all splits.
 *   or_rmat_csr / or_er_csr the build's own generators: pinned by invariants
 *                           (symmetry, no self-loops, sorted unique rows, edge-count
 *                           bounds) and by golden counts of the reference binary
 *                           reading their output as vite files.
 */
#ifndef XXTRIC_ORACLE_H
#define XXTRIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* std::seed_seq{seed}.generate() first word (utils.hpp:37-44 reseeder). */
uint32_t or_reseeder(uint32_t seed);

/* RGG radius of graph.hpp:789-791. */
double or_rgg_radius_ref(int64_t nv);

/* LCG coordinates of utils.hpp:245-430 at one rank: X[0..n), Y[0..n). */
void or_rgg_lcg_coords(int64_t n, uint32_t seed, double* X, double* Y);

/* O(n^2) RGG of graph.hpp:919-948 + CSR of 1289-1341 (one rank, LCG mode).
 * radius <= 0 selects the reference formula.  Returns the number of directed
 * entries ne; fills off[n+1] always and tails[0..ne) when ne <= cap. */
int64_t or_rgg_lcg(int64_t n, uint32_t seed, double radius, int64_t* off,
                   int64_t* tails, int64_t cap, int nthreads);
/* or_rgg_lcg with edge iff sqrt(dx*dx+dy*dy) < radius when strict != 0. */
int64_t or_rgg_lcg_opts(int64_t n, uint32_t seed, double radius, int strict, int64_t* off,
                        int64_t* tails, int64_t cap, int nthreads);

/* LCG coordinates of a P-rank run (utils.hpp:245-430, graph.hpp:836-872): the
 * parallel prefix in wrapping int64 arithmetic, rank r's points at
 * [r*n/P, (r+1)*n/P) with Y in [r/P, (r+1)/P). */
void or_rgg_lcg_coords_ranks(int64_t n, int nprocs, uint32_t seed, double* X, double* Y);

/* O(n^2/P) RGG of a P-rank run (graph.hpp:919-1097, 1289-1341): local pairs,
 * cross pairs with the neighbouring ranks (inner index j > i), one CSR with
 * global ids.  Returns ne, or -1 when P is not a power of two, does not divide
 * n or the radius is not below 1/P. */
int64_t or_rgg_lcg_ranks(int64_t n, int nprocs, uint32_t seed, double radius, int64_t* off,
                         int64_t* tails, int64_t cap, int nthreads);

/* Reference CPU count (xxtric.hpp:60-208, check_edgelist 516-545, pruning
 * 569-588) over vertices vstart, vstart+vstride, ... .  per_vertex (nullable)
 * receives tcount[i] (min-vertex attribution) for processed vertices.
 * nthreads > 1 uses OpenMP schedule(static) as xxtric.hpp:84. */
uint64_t or_count_reference(int64_t nv, const int64_t* off, const int64_t* tails,
                            uint64_t* per_vertex, int nthreads, int64_t vstart,
                            int64_t vstride);
uint64_t or_count_reference32(int64_t nv, const int64_t* off, const int32_t* tails,
                              uint64_t* per_vertex, int nthreads, int64_t vstart,
                              int64_t vstride);

/* or_count_reference32 restricted to a vertex list: per_vertex[k] = tcount of
 * vertices[k] (xxtric.hpp:160-206). */
/* Per-vertex counts (smallest-vertex credit) of the listed vertices by marking
 * N(v) and scanning N(u) above u for each neighbour u > v (rows ascending):
 * the check for hubs, where the reference loop is too slow. */
void or_vertex_min_credit32(int64_t nv, const int64_t* off, const int32_t* tails,
                            const int64_t* vertices, int64_t nlist, uint64_t* out, int nthreads);
void or_count_reference32_list(int64_t nv, const int64_t* off, const int32_t* tails,
                               const int64_t* vertices, int64_t nlist, uint64_t* per_vertex,
                               int nthreads);

/* Time-bounded sample of or_count_reference32 for CPU baselines (vertices in
 * id order, abandoned at the deadline); returns vertices completed. */
int64_t or_count_reference32_budget(int64_t nv, const int64_t* off, const int32_t* tails,
                                    int nthreads, double seconds, uint64_t* triangles,
                                    int64_t* edges_done);

/* Independent exact counter: (degree, id) orientation + sorted merge,
 * per-vertex credit to the smallest id of each triangle. */
uint64_t or_count_degree_ordered(int64_t nv, const int64_t* off, const int32_t* tails,
                                 uint64_t* per_vertex, int nthreads);

/* vite binary edge list (graph.hpp:294-395): int64 nv, ne, (nv+1) offsets,
 * ne x {int64 tail, double weight};  layout32: int32 / float (deftypes.hpp:40-45). */
int or_vite_write(const char* path, int64_t nv, const int64_t* off,
                  const int64_t* tails, int layout32);
/* Reads the header only when off == NULL; returns 0 on success. */
int or_vite_read(const char* path, int layout32, int64_t* nv, int64_t* ne,
                 int64_t* off, int64_t* tails);

/* ---- this build's own seeded generators (host twins of the device ones) ---- */
uint64_t or_mix64(uint64_t z);
void or_rmat_thresholds(double a, double b, double c, uint32_t* t);
/* Raw R-MAT tuples [start, start+count) (before cleaning). */
void or_rmat_tuples(int scale, uint64_t seed, const uint32_t* thr, int permute,
                    int64_t start, int64_t count, uint32_t* src, uint32_t* dst);
/* Cleaned symmetric CSR: self-loops dropped, duplicates merged, rows sorted. */
int64_t or_rmat_csr(int scale, int edgefactor, uint64_t seed, const uint32_t* thr,
                    int permute, int64_t* off, int32_t* tails, int64_t cap);
/* G(n,p) with p given as a 32.32 fixed-point threshold: edge {i<j} iff
 * (mix64(key + i*n + j) >> 32) < thr64. */
int64_t or_er_csr(int64_t n, uint64_t seed, uint64_t thr64, int64_t* off,
                  int32_t* tails, int64_t cap);

/* Batch graph parameters of graph g (log-uniform n, RGG even / ER odd). */
void or_batch_params(uint64_t base_seed, int64_t g, int64_t nmin, int64_t nmax,
                     int64_t* n, int* kind, uint64_t* gseed, uint64_t* er_thr64);

/* ---- vertex partitions for a split over ranks (BinaryEdgeList::read_balanced) ----
 * bins[0..nparts] (zeroed here): rank r owns vertices [bins[r], bins[r+1]).
 * Return 0, or -1 where the reference's table walk would read outside its
 * table (it is undefined there) or nparts > nv - 1 for the dynamic program
 * (the reference's bins are not a partition there). */
/* -b: greedy fill up to ne / nparts entries per rank, graph.hpp:556-602
 * (default build; ne = off[nv] - off[0], the vite header's ne). */
int or_partition_greedy(int64_t nv, const int64_t* off, int64_t nparts, int64_t* bins);
/* -d: linear_partitioning_num_edges (graph.hpp:397-419): Skiena's dynamic
 * program of utils.hpp:46-135 on seq[i] = off[i], i < nv (the file's first nv
 * offsets, as the reference reads them), literally O(nv^2 * nparts). */
int or_partition_linear(int64_t nv, const int64_t* off, int64_t nparts, int64_t* bins);

#ifdef __cplusplus
}
#endif
#endif
