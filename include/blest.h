/*
 * blest.h — C ABI of the MI355X-native BVSS bit-matrix traversal library (libblest.so).
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns a
 * blest_status (never exits the process) and keeps no global mutable state: the
 * reference's globals JACKARD_ON / WINDOW_SIZE / FULL_PADDING (Common.cuh:32-35)
 * become explicit arguments.  Vertex IDs are uint32; CSC column j lists the rows i
 * of the edges j -> i (DataStructures/CSC.cuh:118-131).
 *
 * Ownership: the library owns every handle it returns (free with the matching
 * *_free); the caller owns every output array it passes in.  Host handles are
 * immutable after construction except through blest_csc_reorder.  A device graph
 * may be queried from several host threads at once when each thread passes its own
 * blest_query (per-query device state); queries sharing a blest_query serialise.
 */
#ifndef BLEST_H
#define BLEST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum blest_status {
    BLEST_OK = 0,
    BLEST_ERR_INVALID_ARGUMENT = 1, /* bad size, null pointer, vertex out of range   */
    BLEST_ERR_SLICE_SIZE = 2,       /* BVSS.cuh:103-110 "Invalid slice size/noMasks" */
    BLEST_ERR_PERMUTATION = 3,      /* CSC.cuh:1146-1149 "Permutation check failed"  */
    BLEST_ERR_IO = 4,               /* CSC.cuh:101-104, Benchmark.cuh:70-73          */
    BLEST_ERR_OUT_OF_HOST_MEMORY = 5,
    BLEST_ERR_OUT_OF_DEVICE_MEMORY = 6,
    BLEST_ERR_DEVICE = 7,           /* a HIP runtime call failed (Common.cuh:47-55)  */
    BLEST_ERR_TIMEOUT = 8,          /* a persistent kernel's bounded wait expired    */
    BLEST_ERR_UNSUPPORTED = 9,      /* e.g. CC on a graph whose CSC is not symmetric */
    BLEST_ERR_BUSY = 10             /* serving queue full                            */
} blest_status;

const char* blest_status_str(blest_status s);
/* Text of the last error raised on the calling thread (empty string if none). */
const char* blest_last_error(void);
/* Library version string. */
const char* blest_version(void);

/* ------------------------------------------------------------- CSC graph */
/* Reference: class CSC (DataStructures/CSC.cuh:35-96). */
typedef struct blest_csc blest_csc;

/* Copy a CSC pattern.  Each column is sorted; self-loops are dropped
 * (CSC.cuh:125) and duplicate rows removed (the reference keeps them, which
 * corrupts BVSS masks: BVSS.cuh:246-249).  col_ptrs has n+1 entries. */
blest_status blest_csc_create(uint32_t n, const uint32_t* col_ptrs, const uint32_t* rows,
                              blest_csc** out);
/* Read a MatrixMarket pattern/real file: CSC::CSC(filename, undirected, binary)
 * (CSC.cuh:98-185).  Each line "j i [val]" is the edge j -> i (1-based).  The
 * whole file is read; nothing is downloaded. */
blest_status blest_csc_from_mtx(const char* path, int undirected, int binary, blest_csc** out);
/* Binary cache [u32 n][u32 nnz][u8 isSocial][u32 colPtrs[n+1]][u32 rows[nnz]]
 * (CSC.cuh:225-266). */
blest_status blest_csc_load_binary(const char* path, blest_csc** out);
blest_status blest_csc_save_binary(const blest_csc* g, const char* path);

/* Seeded synthetic generators (counter-based RNG; identical output for a seed on
 * every machine and thread count).  All produce symmetric graphs without
 * self-loops or duplicates, each column sorted. */
/* R-MAT: 2^scale vertices, edge_factor*2^scale samples with quadrant
 * probabilities (a, b, c, 1-a-b-c); vertex IDs permuted by a seeded
 * Fisher-Yates shuffle when permute != 0. */
blest_status blest_csc_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                                     double c, uint64_t seed, int permute, blest_csc** out);
/* side x side 4-neighbour grid, id = r*side + c; each right/down edge kept with
 * probability keep; round(shortcut_frac * side^2) uniform shortcut edges u != v. */
blest_status blest_csc_generate_grid(uint32_t side, double keep, double shortcut_frac,
                                     uint64_t seed, blest_csc** out);
/* Erdos-Renyi G(n, m): m = round(n*avg_degree/2) uniform undirected pairs u != v
 * (duplicates merged). */
blest_status blest_csc_generate_er(uint32_t n, double avg_degree, uint64_t seed, blest_csc** out);

void blest_csc_free(blest_csc* g);
blest_status blest_csc_info(const blest_csc* g, uint32_t* n, uint64_t* nnz, int* is_social);
/* Copy out the pattern: col_ptrs[n+1], rows[nnz]. */
blest_status blest_csc_export(const blest_csc* g, uint32_t* col_ptrs, uint32_t* rows);
/* CSC::isSocialNetwork (CSC.cuh:176,187-217; Common.cuh:100-250). */
blest_status blest_csc_is_social(const blest_csc* g, int* is_social);
/* CSC::checkSymmetry (CSC.cuh:410-435). */
blest_status blest_csc_is_symmetric(const blest_csc* g, int* symmetric);
/* CSC metrics (CSC.cuh:268-333): max/avg bandwidth, max/avg profile, max/avg degree. */
typedef struct blest_csc_metrics {
    uint32_t max_bandwidth, max_profile, max_degree, _pad;
    double avg_bandwidth, avg_profile, avg_degree;
} blest_csc_metrics;
blest_status blest_csc_metrics_get(const blest_csc* g, blest_csc_metrics* m);
/* Seeded roots: `count` distinct vertices of degree >= 1 drawn uniformly. */
blest_status blest_csc_pick_roots(const blest_csc* g, uint32_t count, uint64_t seed,
                                  uint32_t* roots);

/* ------------------------------------------------------------ reordering */
/* Reference: CSC::reorder and the orderings it can reach (CSC.cuh:358-386,
 * 525-1142, Gorder/Graph.cpp:388-617). */
typedef enum blest_reorder_kind {
    BLEST_REORDER_AUTO = 0,           /* CSC::reorder dispatch: social&jaccard -> JACCARD
                                         (or GORDER_JACCARD when gorder != 0),
                                         social&!jaccard -> NATURAL, else RCM */
    BLEST_REORDER_NATURAL = 1,        /* CSC.cuh:525-533 */
    BLEST_REORDER_RCM = 2,            /* CSC.cuh:994-1142 */
    BLEST_REORDER_JACCARD = 3,        /* CSC.cuh:717-920 jackardWithWindow */
    BLEST_REORDER_GORDER = 4,         /* Gorder/Graph.cpp:388-617 */
    BLEST_REORDER_GORDER_JACCARD = 5, /* CSC.cuh:932-948 */
    BLEST_REORDER_RANDOM = 6,         /* CSC.cuh:535-554 (seeded here) */
    BLEST_REORDER_DEGREE = 7          /* CSC.cuh:968-992 ascending degree */
} blest_reorder_kind;

typedef struct blest_reorder_opts {
    blest_reorder_kind kind;
    uint32_t slice_size; /* 8 */
    uint32_t window;     /* Jaccard / Gorder window (reference -w); multiple of slice_size */
    int jaccard;         /* reference -j, used by AUTO */
    int gorder;          /* AUTO: route social&jaccard to GORDER_JACCARD */
    uint32_t threads;    /* 0 = all hardware threads (Gorder: at most 8); every count gives the same order */
    uint64_t seed;       /* RCM pseudo-peripheral restarts, RANDOM */
} blest_reorder_opts;
void blest_reorder_opts_default(blest_reorder_opts* o);

/* Reorder g in place (CSC::applyPermutation, CSC.cuh:1144-1191) and write the
 * inverse permutation inv_perm[old] = new (n entries) when inv_perm != NULL. */
blest_status blest_csc_reorder(blest_csc* g, const blest_reorder_opts* o, uint32_t* inv_perm);
/* Apply a caller-given inverse permutation (checked to be a bijection,
 * CSC.cuh:1193-1210). */
blest_status blest_csc_apply_permutation(blest_csc* g, const uint32_t* inv_perm);
/* Ordering cache [u32 invPerm[n]] (CSC.cuh:335-356). */
blest_status blest_ordering_save(const char* path, const uint32_t* inv_perm, uint32_t n);
blest_status blest_ordering_load(const char* path, uint32_t* inv_perm, uint32_t n);

/* ----------------------------------------------------------------- BVSS */
/* Reference: class BVSS : BitMatrix (DataStructures/BVSS.cuh:28-595). */
typedef struct blest_bvss blest_bvss;

/* BVSS::BVSS(sliceSize, noMasks) + constructFromCSCMatrix (BVSS.cuh:97-113,185-418).
 * full_padding = the reference's FULL_PADDING (Benchmark.cuh:187-194: !isSocial);
 * pass -1 to derive it from the CSC's social classification. */
blest_status blest_bvss_build(const blest_csc* g, uint32_t slice_size, uint32_t no_masks,
                              int full_padding, blest_bvss** out);
/* The same construction on the GPU (slice size 8 with 4 masks only, else
 * BLEST_ERR_UNSUPPORTED): a segmented radix sort of the (row, column bit) keys of each
 * rset, slices numbered by a scan, vsets laid out by prefix sums and a scatter; the
 * arrays (and the update divergence) are bit-identical to blest_bvss_build's and are
 * returned in host memory.  device: HIP device ordinal.  Reference:
 * BVSS::constructFromCSCMatrix (BVSS.cuh:185-467), which the reference runs on the host. */
blest_status blest_bvss_build_device(const blest_csc* g, uint32_t slice_size, uint32_t no_masks,
                                     int full_padding, int device, blest_bvss** out);
void blest_bvss_free(blest_bvss* b);

typedef struct blest_bvss_info {
    uint32_t n, slice_size, no_masks, no_real_slice_sets, no_virtual_slice_sets;
    int full_padding;
    uint64_t no_slices, no_unpadded_slices, no_padded_slices, edges_check;
    double update_divergence;   /* BVSS.cuh:420-467 */
    double avg_slices_per_set, std_slices_per_set;
    uint32_t min_slices_per_set, max_slices_per_set;
    double compression_ratio;   /* BVSS.cuh:532 */
} blest_bvss_info;
blest_status blest_bvss_info_get(const blest_bvss* b, blest_bvss_info* info);
/* Copy out the five BVSS arrays (BVSS.cuh:87-91): slice_set_ptrs[V+1],
 * virtual_to_real[V], real_ptrs[R+1], row_ids[S], masks[S/no_masks]. */
blest_status blest_bvss_export(const blest_bvss* b, uint64_t* slice_set_ptrs,
                               uint32_t* virtual_to_real, uint32_t* real_ptrs,
                               uint32_t* row_ids, uint32_t* masks);
/* Adopt a BVSS built elsewhere, e.g. by the reference's BVSS::constructFromCSCMatrix:
 * the header values and the five arrays exactly as its accessors return them
 * (DataStructures/BVSS.cuh:54-68: getN, getSliceSize, getNoMasks, getNoRealSliceSets,
 * getNoVirtualSliceSets, getNoSlices, getSliceSetPtrs, getVirtualToReal, getRealPtrs,
 * getRowIds, getMasks, getUpdateDivergence; FULL_PADDING as it was built).  The arrays
 * are copied (the caller keeps ownership) and checked: prefix sums, virtualToReal
 * against realPtrs, R = ceil(n / slice_size), vsets of whole mask words, row IDs below n
 * or the pad row 0xFFFFFFFF with an empty mask (BLEST_ERR_INVALID_ARGUMENT otherwise). */
blest_status blest_bvss_adopt(uint32_t n, uint32_t slice_size, uint32_t no_masks, uint32_t no_real_slice_sets,
                              uint32_t no_virtual_slice_sets, uint64_t no_slices, const uint64_t* slice_set_ptrs,
                              const uint32_t* virtual_to_real, const uint32_t* real_ptrs, const uint32_t* row_ids,
                              const uint32_t* masks, double update_divergence, int full_padding, blest_bvss** out);
/* BVSS binary cache (BVSS.cuh:125-183): header + the five arrays. */
blest_status blest_bvss_save_binary(const blest_bvss* b, const char* path);
blest_status blest_bvss_load_binary(const char* path, blest_bvss** out);
/* Statistics row of the reference CSV (BVSS.cuh:112,469-552), tab separated and
 * "'"-prefixed like fileFlush (Common.cuh:75-79); writes at most cap bytes. */
blest_status blest_bvss_csv_header(char* buf, size_t cap);
blest_status blest_bvss_csv_row(const blest_bvss* b, char* buf, size_t cap);

/* ---------------------------------------------------------- device graph */
typedef struct blest_device_graph blest_device_graph;
typedef struct blest_query blest_query;

typedef struct blest_device_opts {
    int device;              /* HIP device ordinal */
    uint32_t block_size;     /* informational: the level kernels run 512-thread workgroups */
    uint32_t blocks_per_cu;  /* 1..8: frontier segments (= workgroups per level kernel) per CU */
    double alpha;            /* top-down -> bottom-up when frontier edges > unexplored edges / alpha */
    double beta;             /* bottom-up -> top-down when frontier vertices < n / beta */
    int direction;           /* 0 = optimizing, 1 = top-down only, 2 = bottom-up after level 1 */
    int keep_csr;            /* upload the bottom-up CSR (needed for direction != 1) */
    uint32_t timeout_ms;     /* bound on every in-kernel wait */
    int use_graphs;          /* level chunks replayed from a captured HIP graph: 0 = once a
                                query's previous traversal was >= 64 levels deep (default),
                                1 = always, 2 = never.  Queries of a request server never replay. */
    uint32_t td_quads;       /* slice tile of a top-down wave step: quads (4 slices) each lane
                                streams per step, 1, 2 (default) or 4 */
    uint32_t bu_group;       /* bottom-up staging depth: bitmap words a wave probes together, 4
                                or 8 (default) */
    int32_t fuse_quads;      /* single-source BFS: a top-down level whose frontier's slices fill
                                at most this many quads (4 slices) runs inside one workgroup,
                                and so do the small levels after it, with no kernel boundary
                                between them; 0 = default (2048), < 0 = never, at most 65536 */
    int32_t bu_probe;        /* bottom-up: each vertex first tests its highest-degree
                                in-neighbour (an n-entry array built at creation), then its
                                list; 0 = default (on), < 0 = off (list order only) */
} blest_device_opts;
void blest_device_opts_default(blest_device_opts* o);

/* Upload a BVSS (and the bottom-up CSR derived from g, the CSR of
 * BVSSBFSKernel::hostCode, BFS/BVSSBFSKernel.cuh:1385-1401) once; queries reuse it.
 * g may be NULL when opts->keep_csr == 0.  perm_old_to_new (nullable) is the
 * inverse permutation the CSC was reordered with: queries then take and return
 * vertices in the original numbering (Benchmark.cuh:237-242,274-282). */
blest_status blest_device_graph_create(const blest_bvss* b, const blest_csc* g,
                                       const uint32_t* inv_perm, const blest_device_opts* o,
                                       blest_device_graph** out);
/* Same, with the bottom-up CSR given as the reference's BVSS keeps it: BVSS::getCSR()
 * (BVSS.cuh:54,187), the transpose of the CSC (column i lists the in-neighbours of i),
 * wrapped with blest_csc_create.  Lets a plugin that holds only the reference BitMatrix
 * upload it without the CSC it was built from. */
blest_status blest_device_graph_create_csr(const blest_bvss* b, const blest_csc* csr,
                                           const uint32_t* inv_perm, const blest_device_opts* o,
                                           blest_device_graph** out);
void blest_device_graph_free(blest_device_graph* dg);
typedef struct blest_device_graph_info {
    uint32_t n, rsets, vsets, grid_blocks, block_size, cus, symmetric, has_csr;
    uint64_t slices, nnz, device_bytes;
    int32_t device;
    uint64_t stream_quads;   /* device slice stream: real slices packed 4 per 20-byte quad */
} blest_device_graph_info;
blest_status blest_device_graph_info_get(const blest_device_graph* dg, blest_device_graph_info* i);

/* Change the direction-switch thresholds (alpha, beta, direction) of a device graph;
 * the layout fields of its opts are fixed at creation.  Not concurrently with a
 * traversal on the graph. */
blest_status blest_device_graph_set_thresholds(blest_device_graph* dg, double alpha, double beta, int direction);

/* Change the tile knobs (td_quads, bu_group) of a device graph; same rules as
 * blest_device_graph_set_thresholds. */
blest_status blest_device_graph_set_tiles(blest_device_graph* dg, uint32_t td_quads, uint32_t bu_group);
/* Change fuse_quads of a device graph (same rules as in blest_device_opts). */
blest_status blest_device_graph_set_fuse(blest_device_graph* dg, int32_t fuse_quads);

/* Autotuner (no reference counterpart; the reference hard-codes its launch shape,
 * BFS/BVSSBFSKernel.cuh:1311-1358): coordinate descent over blocks_per_cu (1, 2, 4, 8),
 * then td_quads (1, 2, 4), bu_group (4, 8), alpha, beta, then fuse_quads (512 to 4096;
 * a base fuse_quads < 0 stays off), each
 * candidate timed as the mean HIP-event time of single-source BFS from `roots`
 * (original IDs) after one warm-up pass; a candidate replaces the best so far only when
 * it is more than 1% faster, so ties between candidates within timing noise keep the
 * earlier (default-side) choice.  base (nullable) supplies the other fields;
 * best receives the fastest opts, best_ms its mean time, trials the candidates run. */
blest_status blest_autotune_bfs(const blest_bvss* b, const blest_csc* g, const uint32_t* perm_old_to_new,
                                const uint32_t* roots, uint32_t nroots, const blest_device_opts* base,
                                blest_device_opts* best, double* best_ms, uint32_t* trials);

/* Per-query device state (levels, bitmaps, frontier lists).  max_sources bounds
 * blest_mbfs batches issued through this query. */
blest_status blest_query_create(blest_device_graph* dg, uint32_t max_sources, blest_query** out);
void blest_query_free(blest_query* q);

typedef struct blest_bfs_stats {
    double kernel_ms;          /* launch + completion, measured with HIP events */
    uint32_t total_levels;     /* max level + 1 (BFS/BVSSBFSKernel.cuh:1551-1561) */
    uint32_t no_visited;
    uint32_t td_levels, bu_levels;
    uint64_t edges_traversed_component; /* 1/2 * sum of degrees of visited vertices */
    uint64_t bottom_up_mask;            /* bit l-1 set: level l ran bottom-up (l <= 64) */
} blest_bfs_stats;

/* BFSKernel::run(src) (BFS/BFSKernel.cuh:53-56): single-source BFS.  levels[n]
 * (nullable) receives hop distances, -1 = unreachable.  Every result array of this
 * header may be host memory (pageable or from blest_host_alloc) or device memory of
 * the graph's device; the copy is ordered on the query's stream and complete on return. */
blest_status blest_bfs(blest_query* q, uint32_t src, int32_t* levels, blest_bfs_stats* st);
/* Same traversal without the device-to-host copy: the levels stay in the query's
 * device buffer (blest_query_levels_device).  stream: hipStream_t or NULL.  The
 * traversal runs on the query's own stream; a non-null caller stream is joined both
 * ways (the traversal starts after the work already queued on it, and work queued on
 * it afterwards starts after the traversal).  blest_query_sync waits for the
 * traversal and fills st.  Both take the query's lock like every other call on it. */
blest_status blest_bfs_async(blest_query* q, uint32_t src, void* stream);
blest_status blest_query_sync(blest_query* q, blest_bfs_stats* st);
/* Device pointer (int32*) of the query's level array (n * max_sources entries). */
void* blest_query_levels_device(blest_query* q);

/* Single-source BFS that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code that must not call hipMalloc, such as
 * tight serving loops or harnesses that forbid it).  Same traversal and result as
 * blest_bfs, without a blest_query: the traversal state lives in a 64 MiB device array
 * that the library holds on each GPU it runs on until the process ends, and the kernels
 * run on `stream` (hipStream_t, NULL = the null stream).  levels: device memory of the
 * graph's device, n int32 entries in original IDs (-1 = unreachable).  A graph whose
 * state exceeds the array (about 4.2 M vertices and 1.1 M vsets, e.g. R-MAT scale 22,
 * fit) is rejected with BLEST_ERR_INVALID_ARGUMENT.  Calls on one GPU share the array
 * and are serialised by a lock held until the traversal has finished; the call returns
 * when the levels are written. */
blest_status blest_bfs_noalloc(const blest_device_graph* dg, uint32_t src, int32_t* levels,
                               blest_bfs_stats* st, void* stream);

/* Multi-source BFS (MBFS/MBFSKernel.cuh:21-53; the reference kernel file is
 * empty): k independent traversals in one launch, k <= max_sources of q.
 * levels[k*n] (row r = root r). */
/* Per-level timing (tracing; no reference counterpart beyond its commented-out
 * per-level clock stamps, BFS/BVSSBFSKernel.cuh:56-58): a single-source BFS that
 * records a HIP event after every kernel.  Level L = 1..total_levels - 1 expands
 * frontier L-1 (bottom_up = direction taken) and folds the result into frontier L. */
typedef struct blest_level_trace {
    uint32_t level, bottom_up;
    float expand_ms, update_ms;
} blest_level_trace;
blest_status blest_bfs_trace(blest_query* q, uint32_t src, int32_t* levels, blest_bfs_stats* st,
                             blest_level_trace* trace, uint32_t cap, uint32_t* count);

typedef struct blest_mbfs_opts {
    int algo;            /* 0 = bit-parallel VALU, 1 = i8 MFMA block product */
    uint32_t batch;      /* roots per launch: 64 or 128 */
} blest_mbfs_opts;
blest_status blest_mbfs(blest_query* q, const uint32_t* srcs, uint32_t k, int32_t* levels,
                        const blest_mbfs_opts* o, blest_bfs_stats* st);

/* CCKernel::run() (CC/CCKernel.cuh:51-54): connected components of a symmetric
 * graph.  labels[n] = smallest original vertex ID of the component (the
 * canonical form; the reference's compaction at CC/BVSSCCKernel.cuh:1284-1290
 * is not reproduced).  algo: 0 = hook + compress over BVSS slices,
 * 1 = min-label propagation over BVSS (CC/BVSSCCKernel.cuh:33-276). */
typedef struct blest_cc_stats {
    double kernel_ms;
    uint32_t no_components;
    uint32_t iterations;
} blest_cc_stats;
blest_status blest_cc(blest_query* q, int algo, uint32_t* labels, blest_cc_stats* st);

/* Page-locked host memory (hipHostMalloc): result arrays placed here come back from the
 * device at full link speed instead of through a staging copy. */
blest_status blest_host_alloc(size_t bytes, void** out);
void blest_host_free(void* p);
/* Free and total HBM of a device (hipMemGetInfo). */
blest_status blest_device_memory_info(int device, uint64_t* free_bytes, uint64_t* total_bytes);
/* Device bytes one query needs for batches of up to max_sources roots. */
blest_status blest_query_footprint(const blest_device_graph* dg, uint32_t max_sources, uint64_t* bytes);

/* ---------------------------------------------------------------- serving */
/* A request server over resident device graphs.  Requests (single-root BFS or CC)
 * queue as they arrive; `workers` host threads, each with its own HIP stream, take
 * the oldest pending request and, for BFS, every other pending BFS request on the
 * same graph (up to max_batch), and run them as one multi-source launch while the
 * other workers' launches proceed on their streams (continuous batching).  A batch
 * starts only when its query state fits the HBM byte budget.  No reference
 * counterpart (the reference has no serving mode). */
typedef struct blest_server blest_server;
typedef struct blest_server_opts {
    uint32_t workers;        /* concurrent batches (streams) */
    uint32_t max_batch;      /* roots per multi-source launch (<= 128) */
    uint64_t hbm_budget;     /* bytes for query state; 0 = 80% of free HBM at creation */
    uint32_t capacity;       /* max queued requests (BLEST_ERR_BUSY beyond) */
    uint32_t min_multi;      /* BFS batches smaller than this run as single-source
                                traversals one after another (a multi-source launch
                                costs about as much as 8 single-source ones) */
} blest_server_opts;
void blest_server_opts_default(blest_server_opts* o);
typedef enum blest_request_kind { BLEST_REQ_BFS = 0, BLEST_REQ_CC = 1 } blest_request_kind;
typedef struct blest_request_result {
    uint64_t ticket;
    int32_t status;          /* blest_status of the request */
    uint32_t batch_size;     /* requests that shared its launch */
    double queued_ms;        /* submit -> launch */
    double service_ms;       /* launch -> results copied */
    double kernel_ms;
    uint32_t total_levels, no_visited;
    uint64_t edges_traversed_component;
} blest_request_result;
blest_status blest_server_create(blest_device_graph* const* graphs, uint32_t count, const blest_server_opts* o,
                                 blest_server** out);
/* Queue a request.  out: caller-owned host array of n entries of that graph
 * (int32 levels for BFS, uint32 labels for CC) that must stay valid until the
 * request completes. */
blest_status blest_server_submit(blest_server* s, uint32_t graph, blest_request_kind kind, uint32_t root,
                                 void* out, uint64_t* ticket);
/* Block until the request completes (or timeout_ms passes: BLEST_ERR_TIMEOUT; 0xFFFFFFFF
 * waits without limit), then release the ticket.  Returns the request's own status. */
blest_status blest_server_wait(blest_server* s, uint64_t ticket, uint32_t timeout_ms,
                               blest_request_result* res);
/* Block until every submitted request has completed. */
blest_status blest_server_drain(blest_server* s);
void blest_server_free(blest_server* s);

/* ------------------------------------------------------- batched graphs */
/* Many small independent graphs, one root each, traversed by one persistent
 * launch (one workgroup per graph, bitmaps in LDS). */
typedef struct blest_graph_batch blest_graph_batch;
/* graphs: count CSC handles (each n <= 65536*4). */
blest_status blest_graph_batch_create(const blest_csc* const* graphs, uint32_t count, int device,
                                      blest_graph_batch** out);
void blest_graph_batch_free(blest_graph_batch* gb);
/* roots[count]; levels: concatenation of each graph's level array (sum n). */
blest_status blest_graph_batch_bfs(blest_graph_batch* gb, const uint32_t* roots, int32_t* levels,
                                   double* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif /* BLEST_H */
