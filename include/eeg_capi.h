/*
 * eeg_capi.h - C-ABI boundary of the MI355X-native push-style graph analytics library.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no C++ or torch types.
 * Every entry point returns an eeg_status; none exits the process (the reference calls
 * exit() on any CUDA or argument error, include/gpu_error_check.cuh:8-16,
 * src/argument_parsing.cpp:51-56).  eeg_last_error() gives a thread-local detail string.
 *
 * Numeric conventions follow the reference:
 *   - unreached BFS/SSSP label = EEG_DIST_INFINITY = UINT_MAX-1   (include/globals.hpp:32)
 *   - SSWP: source label = EEG_DIST_INFINITY, unreached = 0      (src/eegraph.cu:1549-1560)
 *   - CC:   label = minimum vertex id of the weakly connected component
 *           (the reference's cc kernels use SSSP arithmetic, src/cc.cu:62-65; this library
 *           implements min-label propagation on the symmetrised edge list instead)
 *   - PR:   float ranks, x = (1-d)/N + d * sum_{u->v} x_u/outdeg(u), x0 = 1/N; mass at
 *           vertices without out-edges is dropped as in the reference operator
 *           (src/pr.cu:171-225 pushes nothing from degree-0 vertices).
 */
#ifndef EEG_CAPI_H
#define EEG_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EEG_DIST_INFINITY 4294967294u

typedef enum eeg_status {
    EEG_OK = 0,
    EEG_ERR_INVALID_ARG = 1,   /* bad argument (null pointer, source >= n, unsupported K, ...) */
    EEG_ERR_OOM = 2,           /* device or host allocation failed */
    EEG_ERR_DEVICE = 3,        /* a HIP runtime call failed */
    EEG_ERR_NOT_CONVERGED = 4, /* max_iters reached before the fixpoint */
    EEG_ERR_REJECTED = 5,      /* admission control refused a request (label memory) */
    EEG_ERR_NO_DEVICE = 6,     /* no usable GPU */
    EEG_ERR_IO = 7,            /* file could not be read / parsed / written */
    EEG_ERR_UNSUPPORTED = 8,   /* feature not available (e.g. energy counter) */
    EEG_ERR_CANCELLED = 9,     /* a serving request was cancelled before it ran */
    EEG_ERR_TIMEOUT = 10       /* waiting for a serving request timed out */
} eeg_status;

/* Algorithms (reference namespaces bfs/sssp/sswp/cc/pr, include/{bfs,sssp,sswp,cc,pr}.cuh). */
typedef enum eeg_algo {
    EEG_BFS = 0,
    EEG_SSSP = 1,
    EEG_SSWP = 2,
    EEG_CC = 3,
    EEG_PR = 4,
    EEG_PPR = 5 /* personalised PageRank, one-hot restart at the source (serving mix) */
} eeg_algo;

/* Variants, same order and meaning of names as the reference enum Variant
 * (include/globals.hpp:36-41, parsed at src/argument_parsing.cpp:94-106).
 *   ASYNC_PUSH_TD : every virtual node relaxes every sweep, labels in place.
 *   ASYNC_PUSH_DD : data-driven, labels in place; a virtual node runs when its vertex label
 *                   improved since that virtual node last pushed (Gauss-Seidel, no lost
 *                   updates between sibling virtual nodes).
 *   SYNC_PUSH_TD  : same sweep as ASYNC_PUSH_TD (the reference differs only in thread order).
 *   SYNC_PUSH_DD  : frontier-synchronous: the current frontier is processed, improved vertices
 *                   form the next frontier (compacted queue, or per-vertex flags).
 * For PR the sync variants run a Jacobi iteration with destination-owned compensated sums
 * (deterministic); the async variants run the same gather updating ranks in place.
 * Personalised PageRank runs the same gather with a one-hot restart at the source. */
typedef enum eeg_variant {
    EEG_ASYNC_PUSH_TD = 0,
    EEG_ASYNC_PUSH_DD = 1,
    EEG_SYNC_PUSH_TD = 2,
    EEG_SYNC_PUSH_DD = 3
} eeg_variant;

typedef enum eeg_frontier_kind {
    EEG_FRONTIER_SORTED = 0,  /* (default) vertex queue in ascending id order per 8192-vertex
                                 block, built from a 1-bit-per-vertex bitmap after every
                                 iteration; eeg_run_device_noalloc runs EEG_FRONTIER_COMPACT
                                 for it (same answers; the bitmap would need memory) */
    EEG_FRONTIER_FLAGS = 1,   /* per-vertex flag scan over every virtual node */
    EEG_FRONTIER_COMPACT = 2  /* vertex queue in the order vertices improve (ballot/prefix
                                 compaction per tile, dedup stamps) */
} eeg_frontier_kind;

typedef struct eeg_graph eeg_graph;   /* opaque: device-resident CSR + Tigr tables */
typedef struct eeg_server eeg_server; /* opaque: request loop over one eeg_graph */

typedef struct eeg_opts {
    uint32_t variant;          /* eeg_variant, default EEG_ASYNC_PUSH_TD like the reference */
    uint32_t K;                /* virtual-node degree bound: 4, 8, 16 or 32 (0 = 8) */
    uint32_t block;            /* workgroup size of the BFS/SSSP/SSWP/CC relaxation kernels:
                                  256, 512 or 1024 threads (0 = 256; the reference launches 512,
                                  src/eegraph.cu:80-81); PageRank and the batched passes use 256 */
    uint32_t frontier;         /* eeg_frontier_kind, used by EEG_SYNC_PUSH_DD */
    uint32_t max_iters;        /* 0 = no limit beyond the algorithm's own bound */
    uint32_t collect_activity; /* fill stats->activity with the per-iteration frontier size */
    uint32_t timing;           /* time every relaxation launch with HIP events */
    uint32_t count_work;       /* count entries/edges/updates on the device (adds atomics) */
    float damping;             /* PR/PPR damping (0 = 0.85) */
    float acc;                 /* unused; kept for the layout (the CLI's --accuracy sets epsilon) */
    double epsilon;            /* PR and PPR: stop when L1(x_k+1 - x_k) < epsilon (0 = 1e-7) */
    uint32_t wgs_per_cu;       /* workgroups per compute unit of the tile kernels' grid
                                  (0 = default: 24 for frontier queues and topology sweeps, 40
                                  for in-place sweeps, 48 for flag scans and batched passes (96
                                  with narrow labels); PageRank gather 48, in place 24) */
    uint32_t energy;           /* read the amd-smi energy counter before and after the run into
                                  stats->energy_j (two counter reads, ~0.25 ms each; 0 = off) */
    uint32_t label_bits;       /* batched SSSP/SSWP (eeg_run_batch, more than 16 sources): label
                                  width of the planes.  0 = narrow labels: 8-bit widest-path
                                  labels when every weight is <= 254; 8-bit distances when every
                                  weight is <= 127 (queries with a distance past 254 are answered
                                  again on 16-bit planes, see stats->narrow_reruns), else 16-bit
                                  distances (a pass whose distances reach 65535 is rerun with
                                  32-bit labels); 32 = 32-bit labels.  Answers are the same. */
    uint32_t reserved[5];
} eeg_opts;

typedef struct eeg_stats {
    uint32_t iterations;        /* sweeps / frontier rounds that did work */
    uint32_t launches;          /* relaxation kernels launched (includes no-op overshoot) */
    double ms_traversal;        /* device events: first relaxation kernel -> converged */
    double ms_total;            /* host wall clock: label init + traversal + copy-out */
    double ms_kernel_sum;       /* sum of per-launch relaxation kernel times (opts.timing) */
    uint32_t kernels_timed;     /* number of launches summed in ms_kernel_sum */
    uint32_t activity_len;      /* entries written to activity */
    uint64_t edges_relaxed;     /* edges examined */
    uint64_t vnodes_processed;  /* virtual nodes that pushed */
    uint64_t vertices_touched;  /* frontier entries / active vertices staged */
    uint64_t label_updates;     /* successful atomic label improvements */
    double bytes_alg;           /* algorithmic HBM bytes of the traversal */
    double pr_residual;         /* PR: final L1 residual (FP64); PPR: residual mass */
    double energy_j;            /* device energy over the run with opts.energy (else 0, or if unavailable) */
    uint32_t* activity;         /* optional caller buffer for per-iteration frontier sizes */
    uint32_t activity_cap;      /* capacity of activity */
    uint32_t label_bits;        /* batched SSSP/SSWP: label width of the passes that answered
                                   (8, 16 or 32, the widest over passes of 256 sources; 0 for
                                   other runs) */
    uint32_t narrow_reruns;     /* batched SSSP on 8-bit distances: queries answered again with
                                   wider labels because some distance passed 254 */
    uint32_t reserved[5];
} eeg_stats;

typedef struct eeg_graph_info {
    uint32_t n;                 /* vertices */
    uint32_t weighted;          /* 1 if per-edge weights were given */
    uint64_t m;                 /* directed edges as given */
    uint32_t max_out_degree;
    uint32_t argmax_out_degree; /* lowest id among the vertices of maximum out-degree */
    uint32_t nonzero_out;       /* vertices with out-degree > 0 */
    uint32_t device;
    uint64_t device_bytes;      /* HBM held by the handle */
} eeg_graph_info;

/* ---------------- status / device ---------------- */
const char* eeg_strerror(eeg_status s);
const char* eeg_last_error(void);
eeg_status eeg_device_count(int* count);

/* ---------------- host-side graph sources ----------------
 * Seeded synthetic generators (the reference has none; its graphs are SNAP downloads).
 * Counter-based: edge i depends only on (seed, i), so any thread count gives the same list.
 * RMAT: m = edge_factor << scale edges; a,b,c as in Chakrabarti et al., d = 1-a-b-c; ids are
 * not permuted; self-loops and duplicates are kept; w uniform in [wmin, wmax] (w may be NULL).
 */
eeg_status eeg_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                        uint64_t seed, uint32_t wmin, uint32_t wmax,
                        uint32_t* src, uint32_t* dst, uint32_t* w, int threads);
/* 2-D grid width x height, 4-neighbour edges in both directions, plus shortcut pairs chosen
 * uniformly at random (both directions).  n_shortcuts = round(frac * undirected grid edges).
 * eeg_gen_grid_size() reports m so the caller can allocate. */
eeg_status eeg_gen_grid_size(uint32_t width, uint32_t height, double shortcut_frac, uint64_t* m);
eeg_status eeg_gen_grid(uint32_t width, uint32_t height, double shortcut_frac, uint64_t seed,
                        uint32_t wmin, uint32_t wmax, uint32_t* src, uint32_t* dst, uint32_t* w,
                        int threads);

/* Edge-list files.  Text (.el .wel .txt .edges .edgelist) follows src/graph.cpp:35-105:
 * lines not starting with a digit are skipped, "src dst [w]" with w defaulting to 1,
 * n = max id + 1 when id 0 occurs, else max id (ids are then shifted down by one, as
 * src/virtual_graph.cpp:5-12 does, but without mutating anything the caller holds).
 * Binary .bcsr/.bwcsr use the Subway converter layout (Subway/tools/converter.cpp:88-93,
 * 157-162): u32 n, u32 m, u32 offsets[n], then m x {u32 dst} or m x {u32 dst, u32 w}.
 * The library allocates the arrays; release them with eeg_edgelist_free. */
typedef struct eeg_edgelist {
    uint32_t n;
    uint32_t has_zero_id; /* 1 if id 0 occurred in a text file (reference Graph::hasZeroID) */
    uint64_t m;
    uint32_t* src;
    uint32_t* dst;
    uint32_t* w;          /* NULL for unweighted binary files */
} eeg_edgelist;
eeg_status eeg_load_edgelist(const char* path, int weighted, eeg_edgelist* out);
eeg_status eeg_save_edgelist(const char* path, const eeg_edgelist* el);
void eeg_edgelist_free(eeg_edgelist* el);

/* ---------------- device graph ---------------- */
/* Copies the edge list to HBM and builds the out-edge CSR there (stable by input order).
 * src/dst/w are host pointers (w may be NULL: unit weights). Replaces the per-call
 * VirtualGraph(graph); MakeGraph(); cudaMalloc/cudaMemcpy sequence (src/eegraph.cu:15-74). */
eeg_status eeg_graph_create(const uint32_t* src, const uint32_t* dst, const uint32_t* w,
                            uint64_t m, uint32_t n, int device, eeg_graph** out);
/* Same, with src/dst/w already in device memory of `device` (not freed by the library). */
eeg_status eeg_graph_create_device(const uint32_t* d_src, const uint32_t* d_dst,
                                   const uint32_t* d_w, uint64_t m, uint32_t n, int device,
                                   eeg_graph** out);
/* Generates the RMAT list of eeg_gen_rmat directly in HBM (identical edges) and builds it. */
eeg_status eeg_graph_create_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                                 double c, uint64_t seed, uint32_t wmin, uint32_t wmax,
                                 int device, eeg_graph** out);
eeg_status eeg_graph_destroy(eeg_graph* g);
eeg_status eeg_graph_get_info(const eeg_graph* g, eeg_graph_info* info);

/* Builds derived device tables ahead of time so traversals do not pay for them:
 * what bit 0: Tigr tables for K; bit 1: symmetrised CSR (CC); bit 2: in-edge CSR (PR). */
#define EEG_PREP_TIGR 1u
#define EEG_PREP_SYM 2u
#define EEG_PREP_TRANSPOSE 4u
eeg_status eeg_graph_prepare(eeg_graph* g, uint32_t K, uint32_t what);

/* Tigr virtual-graph layout of the reference (src/virtual_graph.cpp:36-112), built on the
 * device and copied out, with defects D1/D2 of the reference fixed (every edge placed):
 *   node_pointer[v] = sum_{u<v} (2 d_u + 1)              (u64, reference stores u32)
 *   edge_list[node_pointer[v]] = d_v, edge j of v at +1+2j (dst) and +2+2j (w)
 *   parts = (v, p) pairs for p < ceil(d_v / K), ascending (v, p)   (PartPointer)
 * edge_list has 2m+n slots, parts has 2*num_parts u32.  Any output pointer may be NULL. */
eeg_status eeg_tigr_export(eeg_graph* g, uint32_t K, uint64_t* node_pointer,
                           uint32_t* edge_list, uint32_t* parts, uint64_t* num_parts);
/* Out-edge CSR as held in HBM (row: n+1 u64, col/w: m u32; w ignored if unweighted). */
eeg_status eeg_csr_export(eeg_graph* g, uint64_t* row, uint32_t* col, uint32_t* w);
/* Derived CSRs (built on demand): which = EEG_PREP_SYM (2m edges: every edge both ways,
 * the list CC runs on) or EEG_PREP_TRANSPOSE (in-edges, the list PageRank gathers over).
 * *m_out receives the edge count; row/col may be NULL to query it. */
eeg_status eeg_csr_export_derived(eeg_graph* g, uint32_t which, uint64_t* row, uint32_t* col,
                                  uint64_t* m_out);

/* ---------------- single query ----------------
 * Replaces Result eegraph_<alg>(ArgumentParser&, Graph&) (include/eegraph.cuh:21-39):
 * labels_out receives n u32 labels (BFS/SSSP/SSWP/CC) or n f32 ranks (PR/PPR); it may be
 * NULL to leave the result on the device.  source is ignored for CC and PR. */
eeg_status eeg_run(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts* opts,
                   void* labels_out, eeg_stats* stats);
/* Same with the labels written to device memory (d_labels_out, n elements) on `stream`
 * (a hipStream_t, NULL = library stream). */
eeg_status eeg_run_device(eeg_graph* g, uint32_t algo, uint32_t source, const eeg_opts* opts,
                          void* d_labels_out, void* stream, eeg_stats* stats);

/* ---------------- single query without allocation ----------------
 * No counterpart in the reference.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  The same query as
 * eeg_run_device (labels identical), but it calls no hipMalloc, hipFree or other allocating
 * routine: the traversal state lives in a fixed device array of the library (about 48 MiB per
 * GPU, held until the process ends), so graphs are limited to EEG_NOALLOC_MAX_VERTICES
 * vertices (larger ones are rejected with EEG_ERR_INVALID_ARG).  BFS, SSSP, SSWP and CC in
 * sync_push_dd with compact queues or the topology-driven sweeps (async_push_dd, flag scans,
 * activity and per-launch timing return EEG_ERR_UNSUPPORTED).  The graph's tables must exist
 * beforehand: eeg_graph_prepare(g, K, EEG_PREP_TIGR | EEG_PREP_SYM).  All calls on one GPU
 * share the array: calls that overlap on one GPU race, so the caller serialises them.
 * stats->ms_traversal is host wall time (no events are created). */
#define EEG_NOALLOC_MAX_VERTICES 2097152u
eeg_status eeg_run_device_noalloc(eeg_graph* g, uint32_t algo, uint32_t source,
                                  const eeg_opts* opts, void* d_labels_out, void* stream,
                                  eeg_stats* stats);

/* ---------------- batched queries ----------------
 * nsrc sources in one traversal: BFS runs multi-source BFS with 64-bit visit masks
 * (ceil(nsrc/64) words per vertex, passes of 256); SSSP/SSWP run vertex-major label planes
 * (passes of 256); PPR runs one pull iteration per source.  labels_out is nsrc x n (u32, or
 * f32 for PPR), row q the answer for sources[q], identical to eeg_run(sources[q]). */
eeg_status eeg_run_batch(eeg_graph* g, uint32_t algo, const uint32_t* sources, uint32_t nsrc,
                         const eeg_opts* opts, void* labels_out, eeg_stats* stats);
/* Several graphs in one launch sequence: graphs[i] with sources[i]; labels_out[i] gets n_i
 * labels.  Every graph must live on the same device. */
eeg_status eeg_run_packed(eeg_graph* const* graphs, uint32_t ngraphs, uint32_t algo,
                          const uint32_t* sources, const eeg_opts* opts, void* const* labels_out,
                          eeg_stats* stats);

/* ---------------- autotuner ----------------
 * No counterpart in the reference, which fixes Part_Size = 8 (include/globals.hpp:30) and
 * takes the variant from --variant (src/argument_parsing.cpp:94-106).  Picks the variant and
 * frontier kind, K, the workgroup size and wgs_per_cu for this graph and algorithm by timing
 * short runs from the
 * sample sources (NULL/0 = the max-out-degree vertex; at most 16 are used).  damping, acc,
 * epsilon and max_iters are read from *best and kept; the other fields are overwritten with
 * the winner.  *best_ms receives its median traversal time.  Choices are cached per graph
 * shape (log2 of N, average degree, degree skew and frontier depth) within the process. */
eeg_status eeg_autotune(eeg_graph* g, uint32_t algo, const uint32_t* sample_sources,
                        uint32_t nsample, eeg_opts* best, double* best_ms);

/* ---------------- serving ----------------
 * No counterpart in the reference (one query per process, src/main.cu:101-121).  A request
 * loop over one device-resident graph.  Engine queries (BFS; PPR; SSSP and SSWP with
 * EEG_SERVER_PLANE_ENGINES) take a column of their engine's planes when one is free
 * (admission by label memory) and join the next iteration of the pass already running
 * (continuous batching); the column is given back at convergence, the answer kept in a result
 * buffer until it is read.  Other queries run on `lanes` concurrent single-query lanes whose
 * launches overlap on the device; a lane batches waiting SSSP/SSWP queries (see
 * EEG_SERVER_NO_LANE_BATCH).  Every answer equals eeg_run's for the same query (PPR:
 * within the stated tolerance, the same iteration summed in another fixed order).  The graph
 * must outlive the server; eeg_run calls on the same graph from other threads are allowed. */
typedef struct eeg_server_cfg {
    uint64_t label_budget_bytes; /* HBM for the engines' label planes (0 = min(free/4, 48 GiB));
                                    split evenly over the BFS, SSSP and SSWP engines, each
                                    column costing 4n bytes of labels + 3n/8 bytes of masks */
    uint32_t max_batch;          /* columns (queries in flight) per engine, <= 256 (0 = 256) */
    uint32_t K;                  /* virtual-node degree bound (0 = 8) */
    float damping;               /* PPR damping (0 = 0.85) */
    double ppr_epsilon;          /* PPR stop: L1(x_k+1 - x_k) < ppr_epsilon (0 = 1e-7) */
    uint32_t flags;              /* EEG_SERVER_* below */
    uint32_t lane_wgs_per_cu;    /* resident workgroups per CU of a single-query lane (0 = 16),
                                    so its kernels leave room for other lanes and the engines */
    uint32_t lanes;              /* single-query lanes (0 = 2, <= 64) */
    uint32_t first_iteration;    /* testing: the traversal engines' iteration counter at start (0 = 1);
                                    from 0xf0000000 on an engine admits nothing until it drains,
                                    then resets its columns and restarts the counter at 1 */
    uint32_t ppr_columns;        /* PPR engine width: 8, 16 or 32 restart vertices per gather
                                    pass (0 = 16).  A pass of 8 / 16 costs about 0.3x / 0.6x
                                    one of 32: narrower engines answer sooner, wider ones serve
                                    more PPR queries per second when many are in flight */
    uint32_t reserved[3];
} eeg_server_cfg;
/* Answers are not kept: columns are given back as soon as a query converges and wait() reports
 * status and times only (labels_out must be NULL).  For throughput runs. */
#define EEG_SERVER_DISCARD_LABELS 1u
/* SSSP and SSWP go through label-plane engines (continuous batching) instead of lanes. */
#define EEG_SERVER_PLANE_ENGINES 2u
/* BFS goes through lanes instead of the multi-source BFS engine. */
#define EEG_SERVER_NO_BFS_ENGINE 4u
/* PPR goes through lanes instead of the PPR engine (rank planes of 8, 16 or 32 restart
 * vertices, one gather pass per iteration for every query in flight). */
#define EEG_SERVER_NO_PPR_ENGINE 8u
/* Lanes answer SSSP and SSWP one query at a time.  By default a lane that takes an SSSP or
 * SSWP query while 31 or more others of the same algorithm wait answers up to 256 of them in
 * one multi-source pass (8-bit label planes when the weights allow, as eeg_run_batch). */
#define EEG_SERVER_NO_LANE_BATCH 16u
typedef struct eeg_query_result {
    uint32_t algo;
    uint32_t source;
    uint32_t batch_size;   /* queries that shared the traversal */
    uint32_t status;       /* eeg_status of the query */
    double t_submit_ms;    /* times on the server clock (ms since server start) */
    double t_start_ms;
    double t_done_ms;
} eeg_query_result;
eeg_status eeg_server_create(eeg_graph* g, const eeg_server_cfg* cfg, eeg_server** out);
/* Queues a query; *ticket identifies it.  Returns EEG_ERR_REJECTED if one query's labels
 * can never fit the budget. */
eeg_status eeg_server_submit(eeg_server* s, uint32_t algo, uint32_t source, uint64_t* ticket);
/* Blocks until the ticket is done (timeout_ms < 0 = forever); copies its labels (n u32 or
 * f32) into labels_out if not NULL, fills res and forgets the ticket (its column is given
 * back).  A traversal query holds its column until it is waited for or cancelled. */
eeg_status eeg_server_wait(eeg_server* s, uint64_t ticket, void* labels_out, double timeout_ms,
                           eeg_query_result* res);
/* Drops a queued query; a running one is forgotten and its column given back when it
 * converges; a finished one is forgotten.  A PPR/PR/CC query already running completes. */
eeg_status eeg_server_cancel(eeg_server* s, uint64_t ticket);
eeg_status eeg_server_destroy(eeg_server* s);

/* ---------------- energy ----------------
 * Device energy accumulator through amd-smi (replaces the NVML power sampler of
 * include/nvmlClass.cuh:121-162,271-355): microjoules since an arbitrary origin. */
eeg_status eeg_energy_read(int device, double* microjoules, uint64_t* timestamp_ns);

#ifdef __cplusplus
}
#endif
#endif /* EEG_CAPI_H */
