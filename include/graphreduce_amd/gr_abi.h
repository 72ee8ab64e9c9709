/*
 * graphreduce_amd C ABI: the drop-in boundary of the Gather-Apply-Scatter
 * engine on one AMD Instinct MI355X (gfx950).
 *
 * The reference drives its GPU engine through the duck-typed C++ class
 * GASEngineGPU<Program> (gpugas.h:1025-1809) whose Program parameter carries
 * the vertex program (bfs.cu:31-80, sssp.cu:25-66, pagerank.cu:27-73,
 * connected_component.cu:25-63).  Templates cannot cross a C ABI, so each
 * Program is instantiated inside the HIP library and selected here by a
 * program id; the C++ header GASEngineGPU.h rebuilds the reference's class
 * on top of these functions.
 *
 * Conventions
 *  - Every function returns GR_OK (0) or a GR_ERR_* code; gr_last_error()
 *    gives the message of the last failure on the calling thread.  (The
 *    reference prints and abort()s, gpugas.h:158-165.)
 *  - Vertex ids, offsets and edge counts are int32, as in the reference
 *    (Int = int32_t, gpugas.h:1026).
 *  - Host pointers unless the name or argument says "device".
 *  - Calls on one engine are not re-entrant; separate engines are
 *    independent and may share one gr_graph.
 */
#ifndef GRAPHREDUCE_AMD_GR_ABI_H
#define GRAPHREDUCE_AMD_GR_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_OK 0
#define GR_ERR_INVALID 1     /* bad argument or size mismatch */
#define GR_ERR_HIP 2         /* HIP runtime failure */
#define GR_ERR_NOMEM 3       /* device or host allocation failed */
#define GR_ERR_STATE 4       /* call out of order (e.g. gather before setGraph) */
#define GR_ERR_UNSUPPORTED 5 /* program/feature not available */
#define GR_ERR_BUDGET 6      /* request exceeds the configured memory budget */

/* Vertex programs compiled into the library (bfs.cu:31, sssp.cu:25,
 * pagerank.cu:27, connected_component.cu:25). */
#define GR_PROGRAM_BFS 1
#define GR_PROGRAM_SSSP 2
#define GR_PROGRAM_PAGERANK 3
#define GR_PROGRAM_CC 4
/* SSSP on congesting links: the gather/apply of sssp.cu with a scatter that
 * adds 1 to every out-edge weight (int32) of an improved vertex whose target
 * is still farther away.  An example of a program with a non-trivial
 * Program::scatter (refgas.h:172-196); its edge data changes during a run. */
#define GR_PROGRAM_SSSP_CONGESTION 5

typedef struct gr_graph gr_graph;
typedef struct gr_engine gr_engine;

/* ---------------------------------------------------------------- runtime */

/* Message of the last failed call on this thread ("" if none). */
const char* gr_last_error(void);

/* Library version string. */
const char* gr_version(void);

/* Number of visible HIP devices (0 on a machine without a GPU; never fails). */
int gr_device_count(int* count);

/* Name, compute-unit count and total HBM bytes of device 0. */
int gr_device_info(char* name, size_t name_len, int* cu_count, size_t* total_bytes);

/* hipDeviceSynchronize on the current device. */
int gr_device_synchronize(void);

/* Plain device buffers, for callers without their own allocator. */
int gr_device_alloc(size_t bytes, void** device_ptr);
int gr_device_free(void* device_ptr);
int gr_memcpy_h2d(void* device_dst, const void* host_src, size_t bytes);
int gr_memcpy_d2h(void* host_dst, const void* device_src, size_t bytes);

/* Streaming device-to-device copy with a plain kernel (16-byte aligned), timed
 * with hipEvents: the reference point for bandwidth and HBM-counter scaling. */
int gr_device_copy(void* device_dst, const void* device_src, size_t bytes, float* ms);

/* The bare random-read pattern of a full gather, timed (average of `repeats`
 * launches after a warm-up): one 4-byte read of a vertex value per CSC slot,
 * value[srcs[e]], with the source ids streamed, over the graph's CSC (of its
 * degree-ordered copy when ordered != 0) — the reads a full gather cannot
 * avoid, without its edge data, reduction or writes. */
int gr_graph_source_read_probe(gr_graph* g, int ordered, int repeats, float* ms);

/* ------------------------------------------------------- host generators */
/* Deterministic synthetic inputs.  All draws are counter-based, so host and
 * device generators give identical edge lists. */

/* RMAT edge list with 2^scale vertices and edge_factor * 2^scale draws,
 * quadrant probabilities a, b, c (d = 1 - a - b - c).  Self-loops and
 * duplicate (src, dst) pairs are removed; the list comes back sorted by
 * (src, dst).  With symmetrize != 0 every arc (u, v) also yields (v, u)
 * before de-duplication.  srcs/dsts must hold capacity entries; *n_edges
 * receives the count.  Pass srcs = dsts = NULL to query the count only. */
int gr_host_rmat_edges(int scale, int edge_factor, double a, double b, double c,
                       uint64_t seed, int symmetrize, int32_t* srcs, int32_t* dsts,
                       int64_t capacity, int64_t* n_edges);

/* side x side 4-neighbour grid (vertex id r*side + c), each undirected edge
 * deleted with probability delete_frac (seeded), both arcs of every kept
 * edge emitted, sorted by (src, dst). */
int gr_host_grid_edges(int side, double delete_frac, uint64_t seed, int32_t* srcs,
                       int32_t* dsts, int64_t capacity, int64_t* n_edges);

/* weights[i] = lo + uniform draw on [0, 2^span_log2) for edge-list position i. */
int gr_host_hash_weights(uint64_t seed, int64_t n, int32_t lo, int span_log2,
                         int32_t* weights);

/* ------------------------------------------------------------------ graph */
/* A device-resident graph: CSC (gather side) and CSR (activation side) plus
 * the permutations from each order back to edge-list order.  Replaces the
 * host builders edgeListToCSR/edgeListToCSC (util.cuh:79-121) used by
 * setGraph (gpugas.h:1346-1373); built on the device with radix sorts. */

/* From a host edge list (any order, duplicates allowed). */
int gr_graph_from_edges(int32_t n_vertices, int32_t n_edges, const int32_t* srcs,
                        const int32_t* dsts, gr_graph** out);

/* RMAT generated on the device; identical to gr_host_rmat_edges. */
int gr_graph_rmat(int scale, int edge_factor, double a, double b, double c,
                  uint64_t seed, int symmetrize, gr_graph** out);

/* Grid generated on the host (gr_host_grid_edges) and uploaded. */
int gr_graph_grid(int side, double delete_frac, uint64_t seed, gr_graph** out);

int gr_graph_destroy(gr_graph* g);
int gr_graph_info(const gr_graph* g, int32_t* n_vertices, int64_t* n_edges);

/* Copies back to the host: CSR offsets[V+1] + dsts[E] (either may be NULL),
 * CSC offsets[V+1] + srcs[E], and the edge list (srcs, dsts) in edge-list
 * order. */
int gr_graph_copy_csr(const gr_graph* g, int32_t* offsets, int32_t* dsts);
int gr_graph_copy_csc(const gr_graph* g, int32_t* offsets, int32_t* srcs);
int gr_graph_copy_edges(const gr_graph* g, int32_t* srcs, int32_t* dsts);

/* Permutations to edge-list order: csr_eid[p] / csc_eid[p] = edge-list index
 * of CSR / CSC position p (either may be NULL). */
int gr_graph_copy_perm(const gr_graph* g, int32_t* csr_eid, int32_t* csc_eid);

/* Vertex of maximum out-degree, lowest id on ties (bfs.cu:175-195). */
int gr_graph_max_out_degree_vertex(const gr_graph* g, int32_t* vertex, int32_t* degree);

/* Device weights for the edge list of g: lo + draw on [0, 2^span_log2) per
 * edge-list position (same values as gr_host_hash_weights). */
int gr_graph_hash_weights_device(const gr_graph* g, uint64_t seed, int32_t lo,
                                 int span_log2, int32_t** device_weights);

/* ----------------------------------------------------------------- engine */
/* One GASEngineGPU<Program> (gpugas.h:1025).  Each call below replaces the
 * member function cited beside it, with refgas.h's semantics
 * (refgas.h:128-229): bulk-synchronous gather over CSC in-edges, apply over
 * the active list, activation of CSR out-neighbours of vertices whose apply
 * returned true, and an ascending active list for the next iteration. */

int gr_engine_create(int program_id, gr_engine** out);
/* Wraps an engine object that code of the caller's instantiated from the
 * library's engine headers for a program of its own
 * (graphreduce_amd::create_engine<Program> in gas_program.hpp, compiled
 * with hipcc against the same headers); the handle then takes every
 * gr_engine_* call and gr_engine_destroy deletes the object.  Replaces the
 * reference's template argument GASEngineGPU<Program> (gpugas.h:1025-1027).
 * Its program id is GR_PROGRAM_CUSTOM. */
#define GR_PROGRAM_CUSTOM 100
int gr_engine_adopt(void* engine_impl, gr_engine** out);
int gr_engine_destroy(gr_engine* e);

/* setGraph (gpugas.h:1314-1319): copies vertexData[n_vertices] (elements of
 * vertex_elem_size bytes, which must equal the program's VertexData) and
 * edge_data[n_edges] (edge-list order; NULL for programs with empty
 * EdgeData) to the device and builds CSR + CSC.  The engine keeps
 * vertex_data to write results back in gr_engine_get_results. */
int gr_engine_set_graph(gr_engine* e, int32_t n_vertices, void* vertex_data,
                        size_t vertex_elem_size, int32_t n_edges, const void* edge_data,
                        size_t edge_elem_size, const int32_t* srcs, const int32_t* dsts);

/* Same as setGraph with a prebuilt device graph (shared, not owned).
 * edge_data is in the graph's edge-list order, on the host or (with
 * edge_data_on_device != 0) on the device. */
int gr_engine_bind_graph(gr_engine* e, gr_graph* g, void* vertex_data,
                         size_t vertex_elem_size, const void* edge_data,
                         size_t edge_elem_size, int edge_data_on_device);

/* setActive (gpugas.h:1524): active = [start, end). */
int gr_engine_set_active(gr_engine* e, int32_t start, int32_t end);
/* countActive (gpugas.h:1548). */
int gr_engine_count_active(gr_engine* e, int32_t* n_active);
/* gather (gpugas.h:1553).  As in refgas.h:143 the flag never skips a
 * non-trivial gather; programs whose gather is constant (BFS) skip it. */
int gr_engine_gather(gr_engine* e, int have_gather);
/* apply (gpugas.h:1621). */
int gr_engine_apply(gr_engine* e);
/* scatterActivate (gpugas.h:1650); Program::scatter runs when
 * have_scatter != 0 (refgas.h:192-196). */
int gr_engine_scatter_activate(gr_engine* e, int have_scatter);
/* nextIter (gpugas.h:1761): builds the next active list; returns its size.
 * Returns as soon as the size is known: the rest of the list's construction
 * may still run on the engine's stream, and every later call on the engine
 * is ordered behind it. */
int gr_engine_next_iter(gr_engine* e, int32_t* n_active);
/* getResults (gpugas.h:1778): copies device vertex data back into the
 * caller's vertexData (the reference's version is compiled out). */
int gr_engine_get_results(gr_engine* e);
/* run (gpugas.h:1792): loop until no vertex is active. */
int gr_engine_run(gr_engine* e, int32_t* iterations);

/* Current iteration number handed to Program::apply (BFS level).  Replaces
 * setIterationCount<true> / g_iterationCountGPU (bfs.cu:27-28, 83-90).  The
 * engine sets it to 0 in setActive and adds 1 in every nextIter. */
int gr_engine_set_iteration_count(gr_engine* e, int32_t iteration);
int gr_engine_get_iteration_count(gr_engine* e, int32_t* iteration);

/* Copies device vertex data into out (n_vertices elements) without touching
 * the bound host array. */
int gr_engine_read_vertex_data(gr_engine* e, void* out, size_t vertex_elem_size);

/* Copies the engine's edge data into out (n_edges elements, the caller's
 * edge-list order).  Programs with a scatter change it during a run: the
 * reference's engine works on the caller's array, so its changes are visible
 * there (refgas.h:69-72, 192-196); GASEngineGPU::getResults copies it back. */
int gr_engine_read_edge_data(gr_engine* e, void* out, size_t edge_elem_size);

/* 1 if the program's scatter may change edge data, 0 if not, -1 for an
 * unknown program id. */
int gr_program_has_scatter(int program_id);

/* Restores the device vertex data to the values given at setGraph/bindGraph
 * with a device-to-device copy (repeated runs without a host round trip;
 * the reference re-runs setGraph for that, bfs.cu:106-108). */
int gr_engine_reset_vertex_data(gr_engine* e);

/* Push/pull switch of scatterActivate: mode 0 = automatic, 1 = push only,
 * 2 = pull only; alpha: pull is chosen when the out-edges of the changed
 * vertices exceed n_edges / alpha. */
int gr_engine_set_direction(gr_engine* e, int mode, double alpha);

/* Gather tile shape: merge-path items per thread (4, 8 or 16; a tile is 256
 * threads).  Chosen by the autotuner; default 8. */
int gr_engine_set_gather_tile(gr_engine* e, int items_per_thread);

/* Internal vertex order (no reference counterpart; default 2).  With 1 the
 * engine runs on a copy of the graph whose vertices are renumbered by
 * descending in+out degree (ties by id), built once per graph on first use
 * and shared by the engines bound to it; every CSC/CSR row keeps its edge
 * order.  The hot vertices' values then share cache lines, which cuts the
 * random traffic of gather and activation.  With 0 the engine runs on the
 * graph as given; with 2 it takes the copy only when some vertex's in+out
 * degree is at least 32 times the mean (an RMAT graph, not a grid or road
 * map, whose own numbering already keeps neighbourhoods local).  Every
 * argument and result stays in the caller's ids (vertex data, setActive
 * ranges, getResults), so results, iteration counts and work counters are
 * those of order 0.  Takes effect at the next setGraph / bindGraph. */
int gr_engine_set_vertex_order(gr_engine* e, int order);

/* Out-of-core mode, the reference's shard streaming (gpugas.h:1398-1503,
 * 1553-1619): max_edges_per_shard > 0 makes the next gr_engine_set_graph keep
 * the edge arrays (CSC sources, CSR targets, edge data) in pinned host memory,
 * cut into vertex ranges of at most that many edges (a vertex with more is a
 * shard of its own); vertex-sized state stays on the device.  Every gather
 * streams the CSC shards that hold active rows, every activation the CSR
 * shards that hold changed vertices, each copy overlapping the previous
 * shard's kernels (two staging buffers, a copy stream).  Results equal the
 * in-memory engine's (FP32 sums may differ in the last place: other tiles).
 * The degree order, the delta/inert gathers and programs whose scatter
 * changes edge data are not used in this mode (the latter: GR_ERR_UNSUPPORTED);
 * gr_engine_bind_graph ignores it.  0 (default): the graph lives in HBM. */
int gr_engine_set_streaming(gr_engine* e, int64_t max_edges_per_shard);

/* Delta-gather programs, adaptive mode: scatterActivate also relaxes the next
 * gather's candidates (one pass over the changed out-edges that marks the
 * next frontier too) when those edges are at most share * E; 0 turns the
 * fusion off, 1 or more fuses every such scatter.  Default 0.4
 * (graphreduce_amd.tune sweeps it per program).  Results are the same either
 * way.  Negative or NaN shares: GR_ERR_INVALID. */
int gr_engine_set_fuse_share(gr_engine* e, double share);

/* Blocked gather (programs whose gather folds int minima of the source value
 * and the edge: SSSP, CC): a gather over at least min_share * E in-edges is
 * done over row blocks of the CSC whose in-edges are listed by source
 * (every row of the graph, one workgroup per block of <= 32768 rows and about
 * E / (2 * compute units) in-edges, at most 2^19, the rows' values kept in
 * LDS), replacing the merge-path tile
 * kernel over the active rows (gpugas.h:714-786, gpugas_kernels.cuh:109-197).
 * The minima are the same; the blocks are built once per graph on first use.
 * Default 0.5; <= 0 turns it off. */
int gr_engine_set_blocked_gather(gr_engine* e, double min_share);

/* Delta gather (no reference counterpart).  For SSSP and CC, once an
 * iteration with every vertex active has run (setActive(0, V)) and every
 * later active set came from scatterActivate, gather may read only the
 * in-edges whose source changed in the previous iteration and treat the
 * others as gatherZero.  Such a source was gathered by the destination after
 * its last change and cannot lower it, so every apply sees the same outcome
 * and value as with the full gather of refgas.h:141-160: results, iteration
 * counts and the gathered_edges counter are unchanged.
 * The delta gather is done either as a pull over the active rows' in-edges
 * that skips unchanged sources, or as a push: the out-edges of the changed
 * vertices lower a per-vertex candidate with atomicMin, read back per active
 * row (the same minimum over the same edges).
 * mode 0: always gather every in-edge; 1 (default): push when under 60 % of
 * the in-edges come from changed sources (gr_engine_set_push_share), else
 * the full gather (the blocked gather, when it applies, skipping the
 * unchanged sources); 2: pull-delta whenever valid; 3: push whenever valid. */
int gr_engine_set_delta_gather(gr_engine* e, int mode);

/* Mode 1 of gr_engine_set_delta_gather pushes when under `share` of the
 * gather's in-edges come from changed sources (default 0.6). */
int gr_engine_set_push_share(gr_engine* e, double share);

/* Work and timing counters, summed since the last reset. */
typedef struct gr_engine_stats {
  int64_t iterations;        /* nextIter calls with a non-empty active set */
  int64_t active_vertices;   /* sum of #A over gathers */
  int64_t gathered_edges;    /* sum of E_in(A) over gathers */
  int64_t changed_vertices;  /* sum of #C (apply returned true) */
  int64_t pushed_edges;      /* sum of E_out(C) */
  int64_t pull_iterations;   /* scatterActivate calls that ran pull mode */
  int64_t gather_launches;   /* gather tile-kernel launches timed */
  double gather_ms;          /* summed duration of the gather tile kernel (pushes: push_ms) */
  int64_t scatter_launches;  /* scatterActivate calls timed (push + pull kernels) */
  double scatter_ms;
  int64_t apply_launches;    /* apply kernel */
  double apply_ms;
  int64_t next_launches;     /* nextIter compaction kernels */
  double next_ms;
  int64_t delta_gathers;     /* gathers that read only sources changed in the
                                previous iteration (see gr_engine_set_delta_gather) */
  int64_t push_gathers;      /* gathers done by pushing the out-edges of the changed
                                (or non-inert) vertices */
  int64_t push_edges;        /* out-edges relaxed by those pushes */
  double push_ms;            /* their summed duration (not in gather_ms) */
  int64_t tile_gather_rows;  /* sum of #A over gathers by the merge-path tile kernel */
  int64_t tile_gather_edges; /* sum of E_in(A) over those gathers */
  int64_t inert_gathers;     /* first gathers that skipped inert sources (SSSP:
                                unreached vertices, weights in [0, 100000]) */
  double inert_ms;           /* their vertex scans (the gather itself is in
                                gather_ms or push_ms) */
  int64_t streamed_bytes;    /* out-of-core mode: edge bytes copied host -> device */
  int64_t streamed_shards;   /* ... in this many shard copies */
  int64_t fused_gathers;     /* push gathers whose relax ran fused with the previous
                                scatterActivate (counted in push_gathers too) */
  int64_t fused_edges;       /* out-edges relaxed by those fused passes (also in
                                push_edges; their time is in push_ms) */
  int64_t blocked_gathers;   /* gathers done by the blocked kernel (counted in
                                tile_gather_rows / tile_gather_edges, timed in
                                gather_ms) */
  int64_t fused_activations; /* of those, the ones that scatterActivate ran in place
                                of the pull, marking the next active set */
  int64_t idle_gather_launches; /* blocked-kernel launches by scatterActivate that
                                stood down on the device (the pull would not run);
                                in gather_launches, their few microseconds in gather_ms */
} gr_engine_stats;

/* Enables hipEvent timing of the engine phases on the engine's stream. */
int gr_engine_enable_timing(gr_engine* e, int on);
int gr_engine_get_stats(gr_engine* e, gr_engine_stats* out);
int gr_engine_reset_stats(gr_engine* e);

/* ------------------------------------------------ PageRank precision form */
/* The recurrence of pagerank.cu:46-66 with every vertex active:
 * r_0 = 0.15, r_{t+1}[v] = 0.15 + 0.85 * sum_{u->v} r_t[u] / outdeg[u], in FP32;
 * the in-edge sums are compensated (TwoSum) when compensated != 0, plain FP32
 * otherwise.  Stops when the FP64 L1 residual sum_v |r_{t+1} - r_t| < tol or
 * after max_iters steps.  The reference's activation-driven variant
 * (tol 0.01 per vertex) is the engine with GR_PROGRAM_PAGERANK. */
typedef struct gr_pagerank_result {
  int32_t iterations;     /* steps run */
  double residual;        /* FP64 L1 residual of the last step */
  double ms;              /* device time of the iteration loop */
  double gather_ms;       /* of which merge-path gather kernels */
  int64_t gathered_edges; /* iterations * E */
} gr_pagerank_result;

/* ranks (host, V floats) and residuals (host, max_iters doubles) may be NULL. */
int gr_pagerank_precise(gr_graph* g, int max_iters, double tol, int compensated, float* ranks,
                        double* residuals, gr_pagerank_result* result);

/* gr_pagerank_precise without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * errors as gr_pagerank_precise (identical ranks and residuals); its buffers
 * come from one static device array per GPU of 80 MiB (83,886,080 bytes),
 * held from the first use until the process ends.  A graph needing more
 * (16 B per vertex + 36 B per 2048 vertices and edges, each buffer rounded
 * up to 256 B; RMAT-22 needs about 69 MB) is refused with GR_ERR_INVALID.
 * Runs on the null stream, creates no stream or event; result->ms is the host
 * clock around the loop and result->gather_ms is 0.  Calls that overlap on
 * one GPU share the array and race: serialise them. */
int gr_pagerank_precise_noalloc(gr_graph* g, int max_iters, double tol, int compensated,
                                float* ranks, double* residuals, gr_pagerank_result* result);

/* ----------------------------------------------- batched multi-source runs */
/* Up to 64 sources share one traversal (no reference counterpart: the
 * reference runs one source per engine, bfs.cu:93-134, sssp.cu:69-111).
 * Results equal independent plain runs of each source. */
typedef struct gr_batch_result {
  int32_t levels; /* BFS levels / SSSP relaxation rounds run */
  double ms;      /* device time from state init to the last kernel */
  int32_t grid;   /* workgroups of the persistent (BFS) or round (SSSP) kernel */
} gr_batch_result;

/* BFS from sources[0..n) with 64-bit visit/frontier masks per vertex: one
 * expand (push or pull) and one update launch per level.  depth_device:
 * device array n*V int32, row s = depths from sources[s] (-1 = unreached),
 * filled by this call.  direction: 0 automatic (pull when the frontier's
 * out-edges exceed E/alpha), 1 push, 2 pull; OR-able options:
 *   GR_MSBFS_PERSISTENT   every level in ONE cooperative launch (grid barriers)
 *   GR_MSBFS_FILTER       pull first checks an in-neighbour's bit in the 1-bit
 *                         frontier map before reading its 8-byte mask (at
 *                         RMAT-24 slower: the hubs first in every row are
 *                         mostly on the frontier)
 *   GR_MSBFS_STREAM_PULL  the streaming pull kernel, currently disabled (the
 *                         call returns GR_ERR_INVALID). */
#define GR_MSBFS_STREAM_PULL 4
#define GR_MSBFS_PERSISTENT 8
#define GR_MSBFS_FILTER 16
int gr_msbfs(gr_graph* g, const int32_t* sources, int n_sources, int32_t* depth_device,
             int direction, double alpha, gr_batch_result* result);

/* SSSP (sssp.cu semantics: distances, unreached = INT_MAX - 100000) from
 * sources[0..n): lane s of a wave relaxes source s; rounds are replayed from a
 * hipGraph until nothing changes.  weights_device: int32 per edge in the
 * graph's edge-list order; dist_device: n*V int32, row s = source s. */
int gr_mssp(gr_graph* g, const int32_t* weights_device, const int32_t* sources, int n_sources,
            int32_t* dist_device, gr_batch_result* result);

/* A batch context: the state of batched runs on one graph (the 64-bit masks,
 * the [V][64] rows, the heavy-row chunk tables, the weights in CSC order and
 * the recorded SSSP rounds), allocated by the first run that needs it and
 * reused by every later run on the context, as an engine binds its graph
 * once.  gr_msbfs / gr_mssp are a context for one run.  weights_device
 * (int32 per edge, edge-list order; read once, at creation) may be NULL for
 * BFS-only use.  The graph must outlive the context; runs on one context are
 * serialised (one caller at a time).  Device memory held at V vertices and
 * E edges: about 282 B per vertex (+ 4 B per edge and 5 B per vertex with
 * weights once SSSP has run). */
typedef struct gr_batch gr_batch;
int gr_batch_create(gr_graph* g, const int32_t* weights_device, gr_batch** out);
/* gr_msbfs on the context (same arguments and results). */
int gr_batch_bfs(gr_batch* b, const int32_t* sources, int n_sources, int32_t* depth_device,
                 int direction, double alpha, gr_batch_result* result);
/* gr_mssp on the context's weights; GR_ERR_INVALID if it has none. */
int gr_batch_sssp(gr_batch* b, const int32_t* sources, int n_sources, int32_t* dist_device,
                  gr_batch_result* result);
int gr_batch_destroy(gr_batch* b);

/* ------------------------------------------------------------------ serving */
/* Continuous-batching query server (no reference counterpart).  Each query
 * kind has a pool of up to 64 resident lanes; a host thread admits queued
 * requests into free lanes as lanes and the memory budget allow, advances
 * every pool by one step per loop (BFS level, SSSP round, PPR iteration; the
 * pools run on separate streams and overlap), and retires finished lanes.
 * Each result equals the query's solo result. */
#define GR_QUERY_BFS 1  /* int32 depth per vertex, -1 unreached */
#define GR_QUERY_SSSP 2 /* int32 distance per vertex, INT_MAX - 100000 unreached */
#define GR_QUERY_PPR 3  /* float rank: r = 0.15 e_seed + 0.85 A^T (r / outdeg), r_0 = e_seed */

typedef struct gr_server gr_server;

typedef struct gr_server_config {
  size_t memory_budget_bytes; /* 0 = unlimited; a lane is admitted only if it fits */
  int max_lanes_per_kind;     /* 0 = 64 */
  int ppr_iterations;         /* 0 = 10 */
} gr_server_config;

typedef struct gr_server_stats {
  int64_t submitted;
  int64_t completed;
  int64_t steps;            /* server loop iterations */
  int32_t max_lanes_in_use; /* over all kinds */
  int64_t peak_bytes;       /* lane bytes admitted at once */
  double p50_ms, p99_ms, max_ms; /* submit -> completion latency */
  double throughput_rps;    /* completed / (last completion - first submit) */
} gr_server_stats;

/* weights_device (int32 per edge, edge-list order) may be NULL when no SSSP
 * query will be submitted.  The graph must outlive the server.  The three
 * pools run on three streams of their own; give the process at least 8 HIP
 * hardware queues (GPU_MAX_HW_QUEUES=8 in the environment before HIP starts;
 * HIP's default is 4), or two pools may share a queue and run their kernels
 * one after another (a BFS level waiting behind a PPR pass). */
int gr_server_create(gr_graph* g, const int32_t* weights_device, const gr_server_config* config,
                     gr_server** out);
int gr_server_destroy(gr_server* s);
/* Thread-safe.  keep_result != 0 keeps the V-element result for gr_server_wait;
 * otherwise only its checksum is kept.  GR_ERR_BUDGET if one lane of this kind
 * can never fit the memory budget. */
int gr_server_submit(gr_server* s, int kind, int32_t vertex, int keep_result, int64_t* request_id);
/* Blocks until the request completed.  result (V elements, may be NULL),
 * checksum (order-independent hash of the result row), latency, steps. */
int gr_server_wait(gr_server* s, int64_t request_id, void* result, uint64_t* checksum,
                   double* latency_ms, int32_t* steps);
int gr_server_get_stats(gr_server* s, gr_server_stats* out);
/* Submits n requests (kinds[i], vertices[i]) with exponential inter-arrival
 * gaps of mean 1/rate (a Poisson stream) from the calling thread. */
int gr_server_load_test(gr_server* s, double rate_per_s, int n_requests, const int32_t* vertices,
                        const int32_t* kinds, int keep_results, int64_t* request_ids);

#ifdef __cplusplus
}
#endif

#endif /* GRAPHREDUCE_AMD_GR_ABI_H */
