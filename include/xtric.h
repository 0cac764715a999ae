/*
 * xtric.h -- C ABI of the MI355X (gfx950) triangle-counting library
 *            (xxtric_amd/lib/libxtric.so).
 *
 * Plain C types only: no HIP, torch or C++ types cross this boundary.  Every
 * function returns an int status (XT_OK == 0); xt_last_error(ctx) holds the
 * message of the last failure on that context.  A context is used by one host
 * thread at a time (calls on one context are serialised by an internal mutex).
 *
 * Counting semantics (all entry points): the input is a simple undirected graph
 * in CSR form -- symmetric adjacency, each row strictly ascending (no duplicate
 * entries); self-loops are allowed and never close a triangle.  Each triangle
 * u < v < w is counted once and, for per-vertex output, credited to u, its
 * smallest vertex id (xxtric.hpp:160-206).  Counts are exact uint64.
 */
#ifndef XTRIC_H
#define XTRIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define XT_OK 0
#define XT_ERR_ARG (-1)       /* bad argument (null pointer, negative size ...) */
#define XT_ERR_HIP (-2)       /* HIP runtime error */
#define XT_ERR_NOMEM (-3)     /* device allocation failed / over budget */
#define XT_ERR_INPUT (-4)     /* adjacency unsorted, duplicated, out of range or asymmetric */
#define XT_ERR_CAPACITY (-5)  /* batch / server capacity exceeded */
#define XT_ERR_NOTFOUND (-6)  /* unknown ticket / slot */
#define XT_ERR_STATE (-7)     /* call not valid in the object's current state */
#define XT_ERR_INTERNAL (-8)  /* unexpected host-side failure (e.g. a thread could not start) */
#define XT_PENDING 1          /* xt_serve_poll: request not finished yet */

/* ---- counting flags ----------------------------------------------------- */
#define XT_ORIENT_DEGREE 0x0u  /* orient u->v iff (deg u, u) < (deg v, v) (default) */
#define XT_ORIENT_ID 0x1u      /* orient u->v iff u < v, the reference's rule */
#define XT_PER_VERTEX 0x2u     /* also produce per-vertex counts (min-id credit) */
#define XT_VALIDATE 0x4u       /* check rows strictly ascending and ids in range */
#define XT_CHECK_SYMMETRY 0x8u /* with XT_VALIDATE: also check (u,v) => (v,u) */
#define XT_CLEAN 0x10u         /* graph input (xt_graph_from_csr / _from_vite): rebuild it on the
                                  device as a simple undirected graph -- rows sorted, duplicate
                                  entries merged, self-loops dropped, missing reverse entries
                                  added -- for files the reference's binary search would misread */

typedef struct xt_ctx xt_ctx;
typedef struct xt_graph xt_graph;

/* Per-phase device times of the last count on a context, from HIP events on
 * the stream the kernels ran on (milliseconds). */
typedef struct xt_timing {
  double orient_ms;   /* degree keys + out-degree count + scans + oriented fill */
  double count_ms;    /* intersection + reduction kernels (the three phases below) */
  double total_ms;    /* first to last event of the count */
  double light_ms;    /* one pass over the input rows: light edges intersected,
                         heavy edges listed under their owners */
  double items_ms;    /* tiles of the owners' item lists (counts, scan, descriptors) */
  double group_ms;    /* persistent group kernel: staged list x streamed lists */
  int64_t ne_plus;    /* oriented edges |E+| */
  int64_t n_heavy;    /* oriented edges intersected by the group kernel */
  int64_t n_tiles;    /* group tiles drawn from the device work queue */
  int64_t n_wave_tiles; /* wave tiles (owners with short out-lists, k_count_waves) */
} xt_timing;

const char* xt_version(void);

/* Context: owns device scratch, one HIP stream, events.  device = HIP ordinal. */
int xt_ctx_create(int device, xt_ctx** out);
int xt_ctx_destroy(xt_ctx* ctx);
const char* xt_last_error(const xt_ctx* ctx);
int xt_last_timing(const xt_ctx* ctx, xt_timing* out);
/* Tuning knobs (set by xxtric_amd.autotune from measurements; 0 keeps the
 * current value): light_cut -- edges whose two out-lists are both at most this
 * long are intersected one per thread, the rest by the group kernel or wave
 * tiles (default 32 for one graph; for batches and the server 8 from 2^22
 * entries on, where wave tiles take the rest, and 64 below; a value set here applies
 * to both, -1 restores the defaults); group_blocks -- persistent group-kernel workgroups per CU (1..4,
 * default 4); group_round -- 16-byte list chunks each group-kernel thread has
 * in flight per streaming round (1, 2 or 4, default 2).  None of them changes
 * a result. */
int xt_ctx_tune(xt_ctx* ctx, int light_cut, int group_blocks, int group_round);
/* Group-tile knobs (0 keeps the current value, -1 restores the default):
 * items_per_tile -- heavy edges of one owner per group tile (64, 128 or 256,
 * default 256); tile_run -- consecutive tiles a group-kernel workgroup draws
 * from the work queue at once (1..64, default 16; an owner's tiles are
 * consecutive, so longer runs keep its table staged more often).  Neither
 * changes a result.  The counterpart of the reference's compile-time tuning
 * macros (counting.cuh:150-176). */
int xt_ctx_tune_tiles(xt_ctx* ctx, int items_per_tile, int tile_run);
/* Wave tiles for plain counts (0 keeps the current value, -1 restores the
 * defaults, -2 turns them off): owners whose out-list holds at most wave_owner
 * ids (1..512) are counted one wave per tile of at most 64 items, with an
 * LDS table per wave of 1 KB (up to 128 ids), 2 KB (256) or 4 KB (512),
 * instead of by the 512-thread group kernel (default 256 for one graph of
 * 2^27 or more entries, 64 for batches and the server from 2^22 entries on,
 * none for smaller inputs, where the extra launch costs more than it saves; a
 * value set here applies at every size).  Plain, per-vertex and owned-range
 * counts all use them.  Does not change a result.
 * No counterpart in the reference, whose launch shapes are compile-time
 * macros (counting.cuh:150-176). */
int xt_ctx_tune_waves(xt_ctx* ctx, int wave_owner);
/* The group kernel's LDS hash of an owner's out-list: 1 (default) the top 11
 * bits of id * 0x9E3779B1, which spreads ids whatever their bit pattern
 * (labels that share low bits, such as an R-MAT graph's without a
 * relabelling); 2 the low 11 bits of the id, one multiply fewer per streamed
 * id, for randomly labelled graphs.  0 keeps, -1 restores the default.  Does
 * not change a result.  No counterpart in the reference. */
int xt_ctx_tune_hash(xt_ctx* ctx, int group_hash);
/* on != 0: count every graph with the layout used for graphs of 2^32 or more
 * entries (64-bit list offsets gathered by the light pass, 16-bit saturated
 * out-degrees for classification) -- for tests and comparisons on graphs of
 * any size.  Changes no result. */
int xt_ctx_force_wide(xt_ctx* ctx, int on);
/* Failure injection for tests: the next n device allocations the context
 * makes for its scratch (counts, batches, servers) fail with XT_ERR_NOMEM as
 * hipMalloc running out would.  0 turns it off. */
int xt_ctx_fail_allocs(xt_ctx* ctx, int n);

/* ---- device memory plumbing ------------------------------------------------
 * The CudaMalloc / CudaMemcpyHtoD / CudaMemcpyDtoH / CudaFree /
 * CudaDeviceSynchronize macros the reference engine uses around its launchers
 * (counting.cuh:40-147, xxtric-gpu.hpp:116-150,266-284), for callers that hold
 * no HIP runtime of their own.  kind: XT_H2D, XT_D2H or XT_D2D. */
#define XT_H2D 1
#define XT_D2H 2
#define XT_D2D 3
#define XT_D2D_KERNEL 4 /* device copy by a plain 16-byte-per-lane copy kernel */
int xt_dev_alloc(xt_ctx* ctx, size_t bytes, void** out);
int xt_dev_free(xt_ctx* ctx, void* p);
int xt_memcpy(xt_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);
int xt_dev_sync(xt_ctx* ctx);
/* No counterpart in the reference: the memory read paths the counting kernels
 * are bound by, measured on this device (HIP events, best of `reps` after one
 * warm-up, into *ms).  kind 0: a streaming read of n 16-byte elements (every
 * byte once, coalesced); kind 1: ngather loads of elem_bytes (2, 4 or 8) from
 * pseudo-random positions of an array of n elements (n a power of two). */
int xt_mem_ceiling(xt_ctx* ctx, int kind, int64_t n, int elem_bytes, int64_t ngather, int reps,
                   double* ms);

/* ---- single-device count of a host CSR ---------------------------------
 * Replaces TriangleCountingSingle(Graph*) + count() (xxtric-gpu.hpp:27-156,
 * 192-270): offsets/tails are GraphElem (int64) host arrays as in
 * Graph::edge_indices_ / Edge::tail_ (graph.hpp:267-269).  Host buffers are
 * borrowed for the duration of the call.  per_vertex (nullable, nv entries)
 * requires XT_PER_VERTEX. */
int xt_count_csr(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* offsets,
                 const int64_t* tails, uint32_t flags, uint64_t* total,
                 uint64_t* per_vertex);

/* ---- lower-FFI drop-ins (device pointers, caller-owned buffers) ---------
 * Same arguments and ownership as the reference launchers; the count is
 * ADDED to *d_local_count (the caller zeroes it, xxtric-gpu.hpp:150).
 * `stream` is passed as the reference passes it: a stream value (hipStream_t)
 * for xt_single_gpu_count and xt_single_gpu_count_masked, the ADDRESS of one
 * (hipStream_t*) for xt_local_count_gpu_inner; NULL = the context's stream. */

/* counting.cuh:259-265 single_gpu_count (counting.cu:537-577): whole-graph count from a device
 * int64 CSR (symmetric, sorted rows), 0 <= nv < 2^31-1 (XT_ERR_ARG otherwise).
 * Returns once the count has been added to *d_local_count (the call reads
 * d_indices[nv] and checks the ids, so it synchronises with `stream`). */
int xt_single_gpu_count(xt_ctx* ctx, int64_t nv, const int64_t* d_indices,
                        const int64_t* d_edge_list, int64_t* d_local_count,
                        void* stream);

/* No counterpart in the reference: xt_single_gpu_count for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, checks and result; the call itself allocates
 * nothing (it still synchronises with `stream`, so it is not for stream
 * capture).  Its scratch is one statically sized 64 MiB device array per GPU,
 * held from the first context created on that GPU until the process ends; a
 * graph needing more (at most 80 bytes per vertex plus 20.7 per entry: e.g.
 * 65,536 vertices and 2.9 M entries fit) is rejected with XT_ERR_CAPACITY
 * before any kernel runs.  All calls on one GPU share that array, so calls
 * that overlapped would race: the library serialises them per device and
 * each returns only after its last kernel has finished. */
int xt_single_gpu_count_noalloc(xt_ctx* ctx, int64_t nv, const int64_t* d_indices,
                                const int64_t* d_edge_list, int64_t* d_local_count,
                                void* stream);

/* counting.cuh:229-245 single_gpu_count_masked (counting.cu:297-359): nedges pairs (e1, m) in
 * d_edge_mask (int64, 2 per edge, built as xxtric-gpu.hpp:91-107); counts
 * |{n in (m, e1) : d_edge_list[n] in N(d_edge_list[m])}| summed over pairs.
 * Asynchronous on `stream` (a hipStream_t value). */
int xt_single_gpu_count_masked(xt_ctx* ctx, int64_t nedges, const int64_t* d_edge_mask,
                               const int64_t* d_indices, const int64_t* d_edge_list,
                               int64_t* d_local_count, void* stream);

/* counting.cuh:314-323 local_count_gpu_inner (counting.cu:640-657): counts
 * |{i in [start,end) : d_edge_list[i] in d_edge_list[jstart..jend)}|.
 * Asynchronous on *(hipStream_t*)stream: like the reference's cudaStream_t*
 * argument (xxtric-gpu.hpp:225 passes &streams[id]), `stream` is the address
 * of a stream, not the stream. */
int xt_local_count_gpu_inner(xt_ctx* ctx, int64_t start, int64_t end, int64_t jstart,
                             int64_t jend, const int64_t* d_edge_list,
                             int64_t* d_local_count, void* stream);

/* ---- device-resident graphs (CSR kept in HBM: int64 offsets, int32 ids) -- */
/* Upload a host CSR; tails_bytes = 4 (int32) or 8 (int64). */
int xt_graph_from_csr(xt_ctx* ctx, int64_t nv, int64_t ne, const int64_t* offsets,
                      const void* tails, int tails_bytes, uint32_t flags, xt_graph** out);
/* Seeded R-MAT generated and cleaned on the device (self-loops dropped,
 * duplicates merged, symmetrised).  permute != 0 relabels vertices by a
 * seeded bijection of [0, 2^scale). */
int xt_graph_rmat(xt_ctx* ctx, int scale, int edgefactor, double a, double b, double c,
                  uint64_t seed, int permute, xt_graph** out);
/* Seeded G(n,p): edge {i<j} iff (mix64(mix64(seed) + i*n + j) >> 32) < thr64. */
int xt_graph_er(xt_ctx* ctx, int64_t n, uint64_t seed, uint64_t thr64, xt_graph** out);
/* Edge list -> resident symmetric CSR, built on the device: pairs (src[i],
 * dst[i]), i < m, in any order and direction, ids in [0, nv) (else
 * XT_ERR_INPUT); self-loops dropped, duplicates merged, both directions kept,
 * rows ascending (a radix sort of (u, v) keys).  The reference builds its CSR
 * on the host (graph.hpp:1289-1341).  1 <= nv < 2^31-1. */
int xt_graph_from_edges(xt_ctx* ctx, int64_t nv, int64_t m, const int64_t* src, const int64_t* dst,
                        xt_graph** out);
/* vite binary file (graph.hpp:294-395; layout32 selects USE_32_BIT_GRAPH)
 * streamed into a resident graph: offsets rebased to 0 (graph.hpp:386-388)
 * and checked monotone, ids checked in [0, nv), read in 4 M-element blocks
 * through two pinned buffers (host memory stays bounded for s26-sized
 * files).  flags: XT_CLEAN (applied first), then XT_VALIDATE /
 * XT_CHECK_SYMMETRY as for xt_graph_from_csr.  The reference assumes sorted,
 * unique rows and checks nothing; without XT_CLEAN, XT_VALIDATE rejects a
 * file whose rows are not strictly ascending (XT_ERR_INPUT). */
int xt_graph_from_vite(xt_ctx* ctx, const char* path, int layout32, uint32_t flags, xt_graph** out);
/* xt_graph_from_edges, xt_graph_count and xt_graph_destroy in one call. */
int xt_count_edges(xt_ctx* ctx, int64_t nv, int64_t m, const int64_t* src, const int64_t* dst, uint32_t flags,
                   uint64_t* total, uint64_t* per_vertex);
int xt_graph_info(const xt_graph* g, int64_t* nv, int64_t* ne);
/* Copy the CSR back to host arrays (offsets nv+1, tails ne). */
int xt_graph_download(xt_ctx* ctx, const xt_graph* g, int64_t* offsets, int32_t* tails);
/* Count a resident graph: orientation + intersection + reduction on device.
 * per_vertex (host, nullable) requires XT_PER_VERTEX. */
int xt_graph_count(xt_ctx* ctx, xt_graph* g, uint32_t flags, uint64_t* total,
                   uint64_t* per_vertex);
/* One rank of a 1-D vertex split (graph.hpp:76-80: contiguous vertex blocks
 * per rank), on one device.  xt_graph_slice keeps the rows of the owned
 * vertices [lo, hi) and of all their neighbours (ghost rows), whole, and
 * empties the rest; ids stay global.  xt_graph_count_owned counts the
 * triangles whose minimum vertex is owned -- the ones the rank owning [lo, hi)
 * credits under the reference's min-vertex rule (xxtric.hpp:165) -- into
 * *owned, and into *local those whose middle vertex is owned too (the rank
 * counts these itself; the others it forwards as (v, w) pairs to the middle
 * vertex's rank, xxtric.hpp:189-204).  *total (nullable) gets every triangle
 * of g, from a second, plain count.  The owned count skips edges whose smaller
 * endpoint is below lo and streams, per edge, only the ids of the shorter
 * list that can complete an owned triangle.  The count is the same on g and on its slice for [lo, hi).  flags
 * selects the orientation; XT_PER_VERTEX is rejected. */
int xt_graph_slice(xt_ctx* ctx, const xt_graph* g, int64_t lo, int64_t hi, xt_graph** out);
int xt_graph_count_owned(xt_ctx* ctx, xt_graph* g, int64_t lo, int64_t hi, uint32_t flags,
                         uint64_t* owned, uint64_t* local, uint64_t* total);
/* Algorithmic byte figure of one count of g under `flags` orientation:
 * 8(n+1) + 4|E+| + 4 * sum over oriented (u,v) of min(d+(u), d+(v)). */
int xt_graph_alg_bytes(xt_ctx* ctx, xt_graph* g, uint32_t flags, uint64_t* bytes,
                       uint64_t* sum_min, int64_t* ne_plus);
/* The same figures split by kernel: heavy edges are those the group kernel
 * intersects (an out-list longer than the light cut). */
typedef struct xt_graph_stats {
  uint64_t alg_bytes;      /* 8(n+1) + 4|E+| + 4 * sum_min */
  uint64_t sum_min;        /* sum over oriented edges of min(d+(u), d+(v)) */
  int64_t ne_plus;         /* |E+| */
  uint64_t sum_min_heavy;  /* sum of min(d+(u), d+(v)) over heavy edges */
  int64_t heavy_edges;     /* number of heavy edges */
  uint64_t group_bytes;    /* 4 * (sum_min_heavy + heavy_edges): streamed ids + item ids */
} xt_graph_stats;
int xt_graph_stats_get(xt_ctx* ctx, xt_graph* g, uint32_t flags, xt_graph_stats* out);
int xt_graph_destroy(xt_ctx* ctx, xt_graph* g);

/* ---- host ingest (C++ in the library; no device work) -------------------- */
/* Reference RGG (graph.hpp:745-1374, one rank): LCG coordinates seeded with
 * reseeder(seed) (utils.hpp:37-44), edge iff sqrt(dx*dx+dy*dy) <= r, r from
 * graph.hpp:789-791 when radius <= 0.  Cell-grid search, same edge set as the
 * reference's all-pairs loop.  Returns ne in *ne; fills tails when cap >= ne. */
int xt_rgg_generate(int64_t nv, uint32_t seed, double radius, int64_t* offsets,
                    int64_t* tails, int64_t cap, int64_t* ne);
/* The RGG the reference builds when run on nprocs MPI ranks (graph.hpp:745-1097,
 * utils.hpp:245-430), all ranks' rows in one CSR with global ids: rank r owns
 * vertices [r*nv/nprocs, (r+1)*nv/nprocs), its LCG stream starts from the
 * reference's parallel prefix (2x2 matrix powers in wrapping int64 arithmetic,
 * as its mat_power leaves them unreduced), its points lie in the strip
 * Y in [r/nprocs, (r+1)/nprocs), and points of neighbouring ranks are joined
 * only when their local indices differ (the reference's cross loops test
 * j > i).  nprocs = 1 is xt_rgg_generate.  XT_ERR_ARG unless nprocs is a
 * power of two dividing nv and the radius is below 1/nprocs (graph.hpp:765-793). */
int xt_rgg_generate_ranks(int64_t nv, int nprocs, uint32_t seed, double radius, int64_t* offsets,
                          int64_t* tails, int64_t cap, int64_t* ne);
double xt_rgg_radius(int64_t nv);
/* xt_rgg_generate_ranks with options (flags, OR-ed):
 *   XT_RGG_STRICT   edge iff sqrt(dx*dx+dy*dy) < r (the reference tests <=,
 *                   graph.hpp:927,982,1014);
 *   XT_RGG_CONNECT  radius <= 0 selects sqrt(2 ln nv / (pi nv)), the connectivity
 *                   radius (xt_rgg_radius_connect), instead of graph.hpp:789-791.
 * flags = 0 is xt_rgg_generate_ranks.  No counterpart in the reference: it
 * builds the random geometric graph that the workload description "edge when
 * Euclidean distance < sqrt(2 ln(n)/(pi n))" names, from the same points. */
#define XT_RGG_STRICT 0x1u
#define XT_RGG_CONNECT 0x2u
int xt_rgg_generate_opts(int64_t nv, int nprocs, uint32_t seed, double radius, uint32_t flags,
                         int64_t* offsets, int64_t* tails, int64_t cap, int64_t* ne);
double xt_rgg_radius_connect(int64_t nv);
/* vite binary file (graph.hpp:294-395); layout32 selects the USE_32_BIT_GRAPH
 * layout.  Header only when offsets == NULL; header and offsets (rebased to
 * start at 0, graph.hpp:386-388) when tails == NULL. */
int xt_vite_read(const char* path, int layout32, int64_t* nv, int64_t* ne,
                 int64_t* offsets, int64_t* tails);
int xt_vite_write(const char* path, int layout32, int64_t nv, const int64_t* offsets,
                  const int64_t* tails);
/* Tab/space separated edge list "src dst [w]" -> symmetric sorted CSR
 * (cugraph/example-inputs format).  Two calls: tails == NULL sizes it. */
int xt_edgelist_to_csr(const char* path, int64_t* nv, int64_t* ne, int64_t* offsets,
                       int64_t* tails);

/* ---- vertex partitions for a split over ranks (host code; no device work) ---
 * Replaces BinaryEdgeList::find_balanced_num_edges (graph.hpp:556-602, flag -b)
 * and linear_partitioning_num_edges (graph.hpp:397-419 with linear_partition,
 * utils.hpp:46-135, flag -d), which read_balanced (graph.hpp:608-640) calls
 * on rank 0 with the file's offsets.  Writes bins[0..nparts]: rank r owns
 * vertices [bins[r], bins[r+1]); the same bins as the reference's.
 *   XT_PARTITION_GREEDY: a vertex joins the current rank while the rank holds
 *     at most (offsets[nv] - offsets[0]) / nparts entries; the last rank takes
 *     the rest.
 *   XT_PARTITION_LINEAR: the smallest possible largest part over the sequence
 *     offsets[0..nv) (the running counts, as the reference reads them), first
 *     split point on ties; O(nv * nparts) time, 4 * nv * nparts bytes.
 *     XT_ERR_ARG when nparts > nv - 1 or where the reference's table walk
 *     would leave its table (its result is undefined there).
 * XT_ERR_INPUT for decreasing or negative offsets; XT_ERR_NOMEM when the
 * linear program's host tables cannot be allocated; nv < 2^31. */
#define XT_PARTITION_GREEDY 0
#define XT_PARTITION_LINEAR 1
int xt_partition_vertices(int64_t nv, const int64_t* offsets, int64_t nparts, int method,
                          int64_t* bins);

/* ---- batching: many small graphs counted by one persistent launch ------- */
typedef struct xt_batch xt_batch;
typedef struct xt_batch_timing {
  double upload_ms;  /* host->device staging of added graphs */
  double run_ms;     /* device time of the batched count (events) */
  int64_t n_graphs, n_vertices, n_edges, n_tiles;
} xt_batch_timing;
#define XT_BATCH_PERSISTENT 0 /* one persistent pipeline over all graphs */
#define XT_BATCH_HIPGRAPH 1   /* same pipeline, captured once and replayed */
#define XT_BATCH_PER_GRAPH 2  /* baseline: one launch sequence per graph */
int xt_batch_create(xt_ctx* ctx, int64_t max_vertices, int64_t max_edges,
                    int64_t max_graphs, xt_batch** out);
/* Append one graph (int32 ids local to the graph, symmetric sorted rows). */
int xt_batch_add(xt_batch* b, int64_t nv, int64_t ne, const int64_t* offsets,
                 const int32_t* tails, int64_t* slot);
/* Append the seeded mix of graphs [first, first+count) (RGG even, ER odd). */
int xt_batch_add_generated(xt_batch* b, uint64_t base_seed, int64_t first, int64_t count,
                           int64_t nmin, int64_t nmax);
int xt_batch_run(xt_batch* b, int mode, uint32_t flags);
int xt_batch_results(xt_batch* b, uint64_t* totals, int64_t n);
int xt_batch_timing_get(const xt_batch* b, xt_batch_timing* out);
/* Algorithmic byte figures of the whole batch under `flags` orientation, as
 * xt_graph_stats_get gives for one graph, summed over its graphs
 * (8(n+1) + 4|E+| + 4 * sum of min(d+(u), d+(v)) each); heavy edges are those
 * whose longer list exceeds the batch light cut.  Orients the batch again on
 * its stream; results of an earlier run stay readable. */
int xt_batch_stats_get(xt_batch* b, uint32_t flags, xt_graph_stats* out);
int xt_batch_clear(xt_batch* b);
int xt_batch_destroy(xt_batch* b);
/* Parameters of batch graph g (log-uniform n in [nmin,nmax], RGG even / ER odd). */
int xt_batch_params(uint64_t base_seed, int64_t g, int64_t nmin, int64_t nmax,
                    int64_t* n, int* kind, uint64_t* gseed, uint64_t* er_thr64);
/* Host CSR of batch graph g (two calls: tails == NULL sizes it). */
int xt_batch_graph_csr(uint64_t base_seed, int64_t g, int64_t nmin, int64_t nmax,
                       int64_t* nv, int64_t* ne, int64_t* offsets, int32_t* tails);

/* ---- serving: continuous batching of arriving graphs -------------------- */
typedef struct xt_server xt_server;
typedef struct xt_serve_config {
  int64_t hbm_budget_bytes; /* admission limit for resident graphs (0 = 8 GiB) */
  int64_t max_batch_graphs; /* graphs folded into one launch (0 = 4096) */
  int64_t max_inflight;     /* batches in flight on the device (0 = 2) */
  int64_t max_queue;        /* pending requests before submit fails (0 = 1<<20) */
  double max_wait_us;       /* batching window after the first queued request */
} xt_serve_config;
typedef struct xt_serve_stats {
  int64_t submitted, completed, rejected, batches;
  double p50_us, p99_us, mean_us, max_us; /* arrival -> result latency */
  double graphs_per_s;                    /* completed / (last done - first arrival) */
  int64_t peak_resident_bytes;
  int64_t isolated_batches;               /* batches holding one large request alone */
  /* Where the worker thread's time went (seconds): admitting and laying out
   * requests, copying them into pinned staging memory (with the copy pool),
   * launching batches (upload + kernels enqueued) and completing them (result
   * read-back, latencies); device_s sums the batches' kernel time on the GPU. */
  double admit_s, stage_s, launch_s, complete_s, device_s;
} xt_serve_stats;
int xt_serve_create(xt_ctx* ctx, const xt_serve_config* cfg, xt_server** out);
/* Non-blocking; the CSR is copied into server-owned host memory before return.
 * arrival_us < 0 stamps the arrival with the server clock. */
int xt_serve_submit(xt_server* s, int64_t nv, int64_t ne, const int64_t* offsets,
                    const int32_t* tails, double arrival_us, uint64_t* ticket);
/* The same without the copy: the server reads offsets/tails in place when it
 * stages the request (checked and copied into pinned memory in one pass, in
 * the server's thread), so the caller keeps them valid and unchanged until
 * xt_serve_poll returns for the ticket or the server is destroyed.  A request
 * that fails the check is reported by xt_serve_poll (XT_ERR_INPUT). */
int xt_serve_submit_borrowed(xt_server* s, int64_t nv, int64_t ne, const int64_t* offsets,
                             const int32_t* tails, double arrival_us, uint64_t* ticket);
/* XT_OK with the count when done, XT_PENDING while queued or running.  A
 * request's own failure (XT_ERR_INPUT: a borrowed request failed its check;
 * XT_ERR_INTERNAL: it could not be staged) retires its ticket like a result.
 * A server-wide error is returned for a request that has not finished, whose
 * ticket stays live: borrowed arrays are released only once
 * xt_serve_ticket_state reports XT_ERR_NOTFOUND. */
int xt_serve_poll(xt_server* s, uint64_t ticket, uint64_t* total, double* latency_us);
#define XT_SERVE_QUEUED 0
#define XT_SERVE_RUNNING 1
#define XT_SERVE_DONE 2
/* XT_SERVE_QUEUED / _RUNNING / _DONE (not yet polled), or XT_ERR_NOTFOUND for a
 * ticket that was never issued or whose result has been reported. */
int xt_serve_ticket_state(xt_server* s, uint64_t ticket);
/* paused != 0 holds admission (queued requests wait, submissions still queue
 * up to max_queue); running batches finish.  For tests and controlled
 * shutdowns. */
int xt_serve_pause(xt_server* s, int paused);
int xt_serve_drain(xt_server* s, double timeout_s);
/* Native request generator for load tests: submits request i (borrowed) of
 * pool graph order[i] at arrival_us[i] on the server clock (waiting until
 * then), drains, polls every ticket and counts results different from
 * expected[graph] (nullable) into *mismatches; latency_us (nullable, nreq)
 * gets each latency, -1 for a request rejected at submit (XT_ERR_CAPACITY,
 * counted in the stats).  The pool arrays stay valid for the call. */
int xt_serve_replay(xt_server* s, int64_t npool, const int64_t* nv, const int64_t* ne,
                    const int64_t* const* offsets, const int32_t* const* tails, const uint64_t* expected,
                    int64_t nreq, const int64_t* order, const double* arrival_us, double timeout_s,
                    int64_t* mismatches, double* latency_us);
int xt_serve_stats_get(xt_server* s, xt_serve_stats* out);
double xt_serve_now_us(const xt_server* s);
int xt_serve_destroy(xt_server* s);

#ifdef __cplusplus
}
#endif
#endif /* XTRIC_H */
