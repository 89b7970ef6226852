/*
 * suminc.h -- C ABI of the MI355X (gfx950) implementation of SumInc's hot path:
 * delta-accumulative PageRank and SSSP over a CSR fragment, optionally compressed
 * into supernodes with entry/exit shortcuts, applied incrementally after an
 * edge-update batch.
 *
 * The reference drives its GPU prototype through C++ functions in namespaces
 * tjn / tjnsssp / tjnpr_precompute / tjnsssp_precompute (cudaCode/ headers (*.cuh)) that it
 * calls from the workers after flattening host CSR arrays and cudaMalloc-ing
 * buffers itself.  Each entry point below replaces one of those calls or one of
 * the flattening/upload blocks; the comment names the reference call site.
 *
 * Conventions
 *   - Every function returns int: SUMINC_OK (0) or a negative SUMINC_E_* code.
 *     suminc_last_error() returns a message for the last failure on a context.
 *     No entry point aborts the process.
 *   - Host pointers are borrowed for the duration of the call only.  Device memory
 *     is owned by the context and released by suminc_ctx_destroy(), except buffers
 *     from suminc_device_alloc(), which the caller frees with suminc_device_free().
 *   - Vertex ids are dense uint32 in [0, n).  Offsets at the boundary are
 *     uint64 (CSR row pointers, n+1 entries).
 *   - SSSP distances are uint32 with SUMINC_INF_U32 (0xFFFFFFFF) for "unreached"
 *     (the reference uses int32 with INT32_MAX, sssp/sssp_ingress.h:103-105).
 *   - One context per host thread; calls on one context are serialised on the
 *     context's own HIP stream.  Functions that return host data synchronise.
 */
#ifndef SUMINC_H_
#define SUMINC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SUMINC_OK 0
#define SUMINC_E_INVALID (-1)      /* bad argument (null pointer, id out of range, ...) */
#define SUMINC_E_HIP (-2)          /* a HIP runtime call failed */
#define SUMINC_E_NOMEM (-3)        /* device allocation failed */
#define SUMINC_E_STATE (-4)        /* call out of order (e.g. no graph uploaded) */
#define SUMINC_E_TOO_LARGE (-5)    /* size beyond what the layout supports */
#define SUMINC_E_NOT_CONVERGED (-6)/* round cap reached before the stop rule held */
#define SUMINC_E_NUMERIC (-7)      /* NaN/Inf detected in the state */

#define SUMINC_INF_U32 0xFFFFFFFFu

/* precision of the PageRank state */
#define SUMINC_PR_FP64 0        /* delta and rank in FP64 */
#define SUMINC_PR_FP32_KAHAN 1  /* FP32 delta, FP32 rank with a Kahan compensation term */

/* edge weight kinds carried by the graph */
#define SUMINC_W_NONE 0
#define SUMINC_W_U32 1
#define SUMINC_W_F32 2

/* update-batch op types: a line "a u v w" or "d u v" of the reference's update
 * file (grape/fragment/inc_fragment_builder.h:38-72) */
#define SUMINC_OP_ADD 0
#define SUMINC_OP_DEL 1

typedef struct suminc_update {
  uint32_t type; /* SUMINC_OP_ADD / SUMINC_OP_DEL */
  uint32_t u, v;
  uint32_t w_u32; /* weight of an add when the graph has SUMINC_W_U32 weights */
  float w_f32;    /* weight of an add when the graph has SUMINC_W_F32 weights */
  uint32_t pad;
} suminc_update;

/* Run parameters; the reference's gflags (examples/analytical_apps/flags.cc)
 * that reach the hot path. */
typedef struct suminc_params {
  double damping;             /* PageRankIngress::d; the reference uses 0.85f (pagerank_ingress.h:23) */
  double termcheck_threshold; /* FLAGS_termcheck_threshold T: stop when sum|delta| < T, a vertex
                                 is active when |delta| > T / N (sum_sync_iter_worker.h:2873-2925) */
  int32_t max_rounds;         /* FLAGS_pr_mr cap on rounds per phase */
  int32_t pr_precision;       /* SUMINC_PR_FP64 or SUMINC_PR_FP32_KAHAN */
  uint32_t sssp_source;       /* FLAGS_sssp_source */
  int32_t pull_threshold_div; /* push->pull switch: pull when active edges > m / div (0: never pull) */
  int32_t shortcut_max_rounds;/* cap on the in-cluster rounds of a shortcut / precompute (the reference's 100,
                                 iter_compressor.h:422) */
  int32_t pull_band_kib;      /* banded pull: a dense PageRank round gathers the contributions of one source
                                 band of at most this many KiB at a time (<= 0: one band, the default) */
  int32_t pull_compact_kib;   /* the plain pull gathers contributions of the vertices with out-edges only,
                                 renamed to dense positions (by decreasing out-degree class, four classes
                                 an octave, id order inside a class; in id order when banded), once the
                                 contributions exceed this many KiB (0: 32768; negative: never) */
  int32_t pull_tile_kib;      /* tiled pull: when the banded and binned pulls are off and the contributions exceed 64 MiB,
                                 a dense round's gather runs one workgroup per tile of destinations (their
                                 sums in LDS) over source windows of this many KiB of contributions, taken
                                 in the same order by every workgroup (0: 1024; negative: never; the
                                 threshold does not apply when set > 0) */
  int32_t pull_bin_kib;       /* binned pull (propagation blocking): when the banded pull is off, a dense round
                                 first copies the contributions named by the in-edges into an edge-ordered
                                 value stream, one source window of this many KiB at a time (every window in
                                 the caches while it is read), then sums the stream by destination block in
                                 LDS; takes precedence over the tiled pull (0: on above 8 MiB of
                                 contributions, 2048 KiB windows; negative: never; > 0: always) */
  int32_t pull_bin_direct;    /* binned pull: the first windows (the hottest sources) that the summing pass
                                 gathers directly instead of through the value stream (negative: the
                                 default, 1) */
  int32_t reserved[2];
} suminc_params;

typedef struct suminc_stats {
  double rounds;              /* rounds executed by the last run/step call sequence */
  double edges_traversed;     /* sum over rounds of adjacency lengths processed (the
                                 reference's f_send_delta_num, pagerank_ingress.h:59-61) */
  double residual;            /* FP64 sum|delta| at stop (PageRank) */
  double kernel_ms;           /* device time of the last run, from HIP events */
  double push_rounds, pull_rounds;
  /* PageRank runs: the FP64 mass balance of the final state, computed on the device.
   * Every push (r += delta, d * delta spread over the out-edges) and every reseed
   * keeps  sum r + sum delta - d * sum_{outdeg(v) > 0} r_v  equal to its initial
   * value (1 - d) (pagerank_ingress.h:27-70; the reference prints the sums in
   * check_result, sum_sync_iter_worker.h:2851-2871).  mass_defect is that sum minus
   * (1 - d): rounding only, for a state that suminc_pr_init started. */
  double mass_defect;
  double mass_total;          /* FP64 sum r + sum delta */
  double reset_vertices;      /* apply_batch: SSSP vertices reset with their dependency subtree */
  double batch_groups;        /* apply_batch: distinct (u, v) pairs named by the ops */
  double touched_rows;        /* apply_batch: out-CSR rows the batch changed */
  double reseed_edges;        /* apply_batch: edges traversed by the PageRank reseed */
  double shortcut_edges;      /* internal edges read by the last shortcut-row build */
  double sketch_rounds;       /* SumInc queries: rounds of the sketch phase */
  double sketch_ms;           /* SumInc PageRank query: device time of the sketch phase */
  double residual_rounds;     /* SumInc PageRank query: rounds of the plain residual phase */
  double precompute_edges;    /* SumInc queries: internal edges read by the cluster precompute */
  double reserved[7];
} suminc_stats;

typedef struct suminc_ctx suminc_ctx;

/* ---- context ---------------------------------------------------------------
 * Replaces the worker's cudaSetDevice/cudaMalloc block and tjn::init's
 * __device__-global aliasing (sum_sync_iter_worker.h:661-691, 778;
 * cudaCode/my_worker.cu:84-122). */
void suminc_default_params(suminc_params* p);
int suminc_ctx_create(int device, const suminc_params* params, suminc_ctx** out);
int suminc_ctx_destroy(suminc_ctx* ctx);
int suminc_set_params(suminc_ctx* ctx, const suminc_params* params);
/* With ctx == NULL: the message of the last failed call without a context on the
 * calling thread (the text-format readers below). */
const char* suminc_last_error(const suminc_ctx* ctx);
int suminc_sync(suminc_ctx* ctx);
/* hipStream_t the context launches on, as an opaque pointer (for event timing). */
int suminc_get_stream(suminc_ctx* ctx, void** stream);

/* ---- graph (CSR fragment in HBM) ---------------------------------------------
 * Replaces the flattening of getOeoffset() into oeoffset_h/cur_oeoff_h/size_oe_h
 * and their upload (sum_sync_iter_worker.h:575-597, 753-776); the device keeps
 * out-CSR and in-CSR (CSC) sorted by (src, dst).  weights: NULL or m values of
 * the kind given. */
int suminc_upload_csr(suminc_ctx* ctx, uint32_t n, uint64_t m, const uint64_t* offsets,
                      const uint32_t* cols, int weight_kind, const void* weights);
/* Seeded R-MAT (a,b,c; d = 1-a-b-c) built directly on the device: edge_factor<<scale
 * draws, self-loops removed, duplicates merged.  No reference counterpart (the
 * reference loads edge files, grape/fragment/loader.h:30-61). */
int suminc_generate_rmat(suminc_ctx* ctx, int scale, int edge_factor, double a, double b,
                         double c, uint64_t seed, int weight_kind, uint32_t wmin, uint32_t wmax);
int suminc_graph_info(suminc_ctx* ctx, uint32_t* n, uint64_t* m, int* weight_kind);
int suminc_download_csr(suminc_ctx* ctx, uint64_t* offsets, uint32_t* cols, void* weights);
/* The in-adjacency (CSC) the device keeps beside the CSR: sources of each vertex's
 * in-edges, sorted, with their weights (the fragment's incoming adjacency,
 * GetIncomingAdjList in immutable_edgecut_fragment.h). */
int suminc_download_csc(suminc_ctx* ctx, uint64_t* offsets, uint32_t* rows, void* weights);

/* ---- update batches ------------------------------------------------------------
 * Seeded batch: n_del deletes of existing edges, then n_ins inserts with uniform
 * endpoints and fresh weights (synthetic input; the reference generates update
 * files offline with expr2/sh/gen_inc2.py). */
int suminc_generate_batch(suminc_ctx* ctx, uint64_t seed, uint32_t batch_id, uint64_t n_del,
                          uint64_t n_ins, uint32_t wmin, uint32_t wmax, suminc_update* ops_out);
/* Apply a batch to the device CSR/CSC with IncFragmentBuilder semantics
 * (inc_fragment_builder.h:29-133: first add of a pair wins, add beats delete, add of
 * an existing pair replaces its weight).  Replaces reloadGraph()'s host rebuild and
 * re-upload (sum_sync_iter_worker.h:214-345, 1527-1678).  Before the topology
 * changes, the PageRank state (if initialised) is reseeded with AmendValue(-1) on
 * the old out-edges and AmendValue(+1) on the new out-edges of every source named
 * by an op (the active-vertex form the reference left commented out at :285-290,
 * :338-342; equal in exact arithmetic to its full sweeps), and the SSSP state (if
 * initialised) is repaired: every vertex whose parent edge was deleted or re-weighted
 * is reset with its dependency subtree (sum_sync_traversal_worker.h:146-261), reset
 * vertices re-pull from surviving in-neighbours and every insert is relaxed. */
int suminc_apply_batch(suminc_ctx* ctx, const suminc_update* ops, uint64_t n_ops);
/* The same with the ops already in device memory of the context's GPU (for
 * pipelines that keep update batches resident in HBM).  The ops are checked on the
 * device before use; no counterpart in the reference, whose batches are parsed
 * from host files (inc_fragment_builder.h:29-133). */
int suminc_apply_batch_device(suminc_ctx* ctx, const suminc_update* dev_ops, uint64_t n_ops);

/* ---- device buffers owned by the caller (hipMalloc / hipFree / copy on the
 * context's device and stream); plumbing for device-resident inputs. */
int suminc_device_alloc(suminc_ctx* ctx, uint64_t bytes, void** dev_ptr);
int suminc_device_free(suminc_ctx* ctx, void* dev_ptr);
/* src may be host memory or device memory of the context's GPU */
int suminc_copy_to_device(suminc_ctx* ctx, void* dev_dst, const void* host_src, uint64_t bytes);
/* device memory of the context's GPU to host memory (synchronises) */
int suminc_copy_from_device(suminc_ctx* ctx, void* host_dst, const void* dev_src, uint64_t bytes);

/* ---- PageRank (plain Ingress path) --------------------------------------------
 * init: delta = (1-d)/N, rank = 0 (PageRankIngress::init_c/init_v,
 * pagerank_ingress.h:27-32; IterateKernel::Init, ingress_app_base.h:141-228). */
int suminc_pr_init(suminc_ctx* ctx);
/* Runs delta-PageRank rounds until sum|delta| < T or max_rounds: replaces
 * tjn::g_function_pr + tjn::deltaSum (sum_sync_iter_worker.h:885-893, 1372;
 * cudaCode/my_worker.cu:124-142, 304-328).  Returns SUMINC_E_NOT_CONVERGED if the
 * cap is hit (state stays valid). */
int suminc_pr_run(suminc_ctx* ctx, suminc_stats* stats);
/* One round; residual_out = FP64 sum|delta| before the round. */
int suminc_pr_step(suminc_ctx* ctx, double* residual_out);
/* Copies rank (rank - compensation for FP32/Kahan) and delta to host as FP64. */
int suminc_pr_download(suminc_ctx* ctx, double* rank, double* delta);
/* Sets the PageRank state from host FP64 arrays (n each). */
int suminc_pr_upload(suminc_ctx* ctx, const double* rank, const double* delta);

/* ---- SSSP (plain Ingress path) ------------------------------------------------
 * init: dist = inf, parent = self; source: dist 0 (SSSPIngress::GetInitValue /
 * GetInitDelta, sssp/sssp_ingress.h:25-41). */
int suminc_sssp_init(suminc_ctx* ctx, uint32_t source);
/* Min-plus frontier rounds to a fixed point: replaces tjnsssp::g_function +
 * setCurSize/getAllSize (sum_sync_traversal_worker.h:1090, 1136-1142;
 * cudaCode/my_ssspworker.cu:117-495) with weighted relaxation (the reference GPU
 * path uses unit weights, my_ssspworker.cu:390). */
int suminc_sssp_run(suminc_ctx* ctx, suminc_stats* stats);
int suminc_sssp_download(suminc_ctx* ctx, uint32_t* dist, uint32_t* parent);
/* One min-plus round from the pending frontier (the source after
 * suminc_sssp_init, the seeds of a batch, or the previous step's output), for a
 * worker that keeps its own loop: replaces one tjnsssp::g_function call with
 * setCurSize/getAllSize (sum_sync_traversal_worker.h:1090, 1136-1142).
 * *frontier_out gets the size of the frontier the round produced, which becomes
 * the pending one (0: fixed point). */
int suminc_sssp_step(suminc_ctx* ctx, uint32_t* frontier_out);
/* The other traversal Ingress apps on the same kernels, state, batch repair and
 * single step (suminc_sssp_run / _step / _download and suminc_apply_batch then act
 * on the chosen app; the supernode SSSP query takes SUMINC_APP_SSSP only):
 *   SUMINC_APP_SSSP  min-plus over the edge weights (sssp_ingress.h:131-175);
 *   SUMINC_APP_BFS   hops from the source, every edge counting 1 whatever its
 *                    weight (bfs/bfs_ingress.h:60-95);
 *   SUMINC_APP_SSWP  widest path: the largest over paths of the smallest edge
 *                    weight (sswp/sswp_ingress.h:60-140); the source holds
 *                    0xFFFFFFFF, an unreached vertex 0 (its identity element).
 * The value's word is dist (SSSP, BFS) or ~width (SSWP), so one 64-bit atomicMin
 * serves all three.  SSSP and BFS keep the smallest-id predecessor among the
 * optimal ones as parent; SSWP changes a vertex only on a strictly wider path (as
 * SetIfGreaterThan does), so its parent is a predecessor whose strict improvement
 * gave the final width (the smallest id among those of one round) and the parents
 * never form a cycle. */
#define SUMINC_APP_SSSP 0
#define SUMINC_APP_BFS 1
#define SUMINC_APP_SSWP 2
int suminc_traversal_init(suminc_ctx* ctx, int app, uint32_t source);

/* PHP, penalized hitting probability (php/php_ingress.h:15-120), on the weighted
 * series rounds: delta_0 = 1 at the source, a vertex u sends d * w_uv / s_u of its
 * delta along each out-edge to v != source, s_u = the sum of u's out-edge weights to
 * vertices other than the source (weight_sum, :30-45; unit weights on an unweighted
 * graph); values[v] (n FP64, may be NULL) = the sum of the deltas v took, until
 * the round's sum |delta| < termcheck_threshold.  The reference's d is 0.80f
 * (php_ingress.h:31). */
int suminc_php_run(suminc_ctx* ctx, uint32_t source, double damping, double* values, suminc_stats* stats);

/* ---- supernode sketch ------------------------------------------------------------
 * Clustering input, acceptance, node types, shortcut precompute and the SumInc
 * PageRank query (IterCompressor::init/run + SumSyncIterWorker::Query in the
 * mirror-free form, mirror_k = 1e8: compressor_base.h:3160-3611, 4107-4182,
 * 4636-4899; iter_compressor.h:65-196, 318-491, 534-785; sum_sync_iter_worker.h:
 * 348-438, 949-1127, 1385-1491). */
typedef struct suminc_sketch_info {
  uint64_t clusters;            /* accepted clusters (supernode groups) not dissolved */
  uint64_t raw_clusters;        /* clusters in the input assignment */
  uint64_t entries;             /* entry vertices = supernodes (Fc >= 0) */
  uint64_t clustered_vertices;
  uint64_t node_type_count[5];  /* Single, OnlyOut, OnlyIn, BothOutIn, Inner (grape/types.h:70-78) */
  uint64_t sketch_edges;        /* ib_e: Singles' out-edges + exits' external edges */
  uint64_t internal_edges;      /* edges inside accepted clusters */
  uint64_t shortcut_slots;      /* sum over entries of their cluster size (dense rows) */
  uint64_t max_cluster;
  /* accepted clusters dissolved after batches: the acceptance rule of the build
   * (internal edges - entries * exits > compress_threshold) is checked again for
   * every cluster a batch touches; one that fails it has its members laid out as
   * Singles from then on (their cluster_of reads SUMINC_INF_U32) */
  uint64_t dissolved_clusters;
  uint64_t reserved[7];
} suminc_sketch_info;
/* Built-in clustering: every vertex whose only neighbour (in or out) is h joins
 * h's cluster (a pair anchors on its smaller id); clusters of fewer than
 * min_node_num members are dropped, larger ones keep the first max_node_num - 1
 * leaves by id.  Stands in for the reference's external Louvain file
 * (sh/louvain_all.sh:24-49, read at compressor_base.h:3175-3236). */
int suminc_cluster_leaves(suminc_ctx* ctx, uint32_t max_node_num, uint32_t min_node_num,
                          uint32_t* cluster_of_out /* n, may be NULL */);
/* Caller-supplied assignment: cluster_of[v] = cluster id (< n) or SUMINC_INF_U32. */
int suminc_set_clusters(suminc_ctx* ctx, const uint32_t* cluster_of);
/* Acceptance + node types + layout (the structure is rebuilt after every batch,
 * acceptance is kept). */
int suminc_sum_build(suminc_ctx* ctx, uint32_t min_node_num, uint32_t max_node_num, double compress_threshold);
int suminc_sum_info(suminc_ctx* ctx, suminc_sketch_info* info);
/* cluster_of: accepted cluster per vertex (SUMINC_INF_U32 for Singles); node_type: u8 per vertex. */
int suminc_sum_download(suminc_ctx* ctx, uint32_t* cluster_of, uint8_t* node_type);
/* PageRank shortcut of one entry: its cluster's members (exits first), reach mass
 * (bound_delta + inner_value) and residual (inner_delta) per member. */
int suminc_sum_pr_shortcut(suminc_ctx* ctx, uint32_t entry, uint32_t cap, uint32_t* members, double* reach,
                           double* residual, uint32_t* count, uint32_t* nexit);
/* The SumInc PageRank query from the current state: precompute of dirty clusters,
 * sketch rounds, flush, correction, plain residual phase.  Replaces
 * tjn::g_function_compr + OutMirrorSyncToMaster + deltaSum and the host correction
 * (sum_sync_iter_worker.h:932-933, 1129-1146, 1385-1491). */
int suminc_sum_pr_run(suminc_ctx* ctx, suminc_stats* stats);
/* SSSP shortcut of one entry: its cluster's members (exits first, count of them),
 * and for each of the first nexit members (the exits) the in-cluster shortest
 * distance from the entry and the in-cluster parent (the entry itself: 0 and
 * SUMINC_INF_U32; unreachable: SUMINC_INF_U32 twice); slots past the exits are
 * SUMINC_INF_U32 (the row keeps exits only).  Replaces the
 * traversal index of TravCompressor::build_trav_index_mirror /
 * fianl_build_trav_index_mirror (grape/fragment/trav_compressor.h:244-431). */
int suminc_sum_sssp_shortcut(suminc_ctx* ctx, uint32_t entry, uint32_t cap, uint32_t* members, uint32_t* dist,
                             uint32_t* parent, uint32_t* count, uint32_t* nexit);
/* The SumInc SSSP query from the current state (after suminc_sssp_init or an
 * update batch): precompute of the clusters holding the source or a frontier
 * member, sketch rounds (Singles and exits relax their sketch edges, entries
 * their shortcut row to the exits), correction of every member.  Distances and
 * parents equal suminc_sssp_run's.  Replaces the compressed branch of
 * SumSyncTraversalWorker::Query (grape/worker/sum_sync_traversal_worker.h:1160-1233)
 * and its correction (:1322-1406). */
int suminc_sum_sssp_run(suminc_ctx* ctx, suminc_stats* stats);

/* ---- the reference compressor's sketch, driven phase by phase -------------------
 * For a worker that keeps the reference's compressor (host) and its phase machine
 * (batch -> compressed rounds -> correction -> Ingress rounds -> update,
 * sum_sync_iter_worker.h:1372-1694): the compressor's products cross the ABI as
 * they are (compressor_base.h:4636-4899, 5974-5987; iter_compressor.h:433-491),
 * mirrors included.  Ids n .. all_node_num-1 are mirrors.  Row pointers have
 * rows + 1 entries; every pointer is host memory, copied during the call.  After
 * a batch changes the graph the sketch must be uploaded again (the reference
 * re-runs sketch2csr_mirror after inc_run). */
typedef struct suminc_sketch_csr {
  uint32_t n;                    /* vertices: must equal the uploaded graph's */
  uint32_t all_node_num;         /* n + mirrors */
  const uint8_t* node_type;      /* n: NodeType 0..6 (get_nodetype_mirror, compressor_base.h:4144-4182) */
  const uint32_t* mirror_master; /* all_node_num - n: the vertex each mirror copies (mirrorid2vid) */
  const uint64_t* is_off;        /* all_node_num rows: is_e_, a supernode source's bound_delta */
  const uint32_t* is_col;        /*   (exits and out-mirrors of its cluster) */
  const float* is_w;
  const uint64_t* ib_off;        /* n rows: ib_e_, a Single's out-edges, an exit's external ones */
  const uint32_t* ib_col;
  const uint64_t* sync_off;      /* all_node_num rows: sync_e_, master -> its in-mirrors, */
  const uint32_t* sync_col;      /*   out-mirror -> its master */
  const uint64_t* iv_off;        /* all_node_num rows: inner_value of the supernode a vertex or */
  const uint32_t* iv_col;        /*   in-mirror is the source of (SuperNodeForIter, super_node.h:8-36) */
  const float* iv_w;
  const uint64_t* id_off;        /* all_node_num rows: its inner_delta */
  const uint32_t* id_col;
  const float* id_w;
  const uint64_t* sub_off;       /* all_node_num rows, optional (NULL: no precompute): the cluster */
  const uint32_t* sub_col;       /*   subgraph (build_subgraph_mirror, compressor_base.h:3786-3856): a
                                  *   member's out-edges inside its cluster plus the out-mirror ids of the
                                  *   edges its cluster's out-mirrors carry; an in-mirror's master's edges
                                  *   into the cluster; empty for Singles and out-mirrors */
} suminc_sketch_csr;
/* Replaces the worker's flattening of is_e_/ib_e_/sync_e_ and node types and their
 * upload with tjn::init (sum_sync_iter_worker.h:573-784).  The sketch accumulators
 * (bound_node_values, spnode_datas) start at zero and persist across
 * suminc_pr_step_phase calls until suminc_pr_correct, a new PageRank state
 * (suminc_pr_init / suminc_pr_upload) or another upload. */
int suminc_upload_sketch(suminc_ctx* ctx, const suminc_sketch_csr* sketch);
#define SUMINC_PHASE_PLAIN 0  /* an Ingress round over the full CSR (suminc_pr_step) */
#define SUMINC_PHASE_SKETCH 1 /* a round over the uploaded sketch by node type, then the out-mirror sync */
/* One round of the given phase; residual_out = the FP64 residual before the round
 * (plain: sum |delta|; sketch: sum |delta| + sum |bound_node_values|).  The sketch
 * round replaces tjn::g_function_compr + tjn::OutMirrorSyncToMaster + deltaSum
 * (sum_sync_iter_worker.h:932-933, 1129-1142; cudaCode/my_worker.cu:235-293,
 * 331-437). */
int suminc_pr_step_phase(suminc_ctx* ctx, int phase, double* residual_out);
/* The precompute that opens the compressed phase (sum_sync_iter_worker.h:387-435;
 * iter_compressor.h:534-560, 714-765): every cluster member's pending delta runs
 * to convergence inside its cluster over the subgraph rows (Jacobi rounds of all
 * clusters at once, until the mass moved in a round is at most termcheck_threshold
 * / n per member or shortcut_max_rounds; what is left stays in delta), taken into
 * rank; the out-mirrors pass what they collected to their masters' delta; then
 * every exit sends the mass that passed through it along its ib_e_ row and, as a
 * master, into its in-mirrors' rows (spnode_datas, bound_node_values).  Needs the
 * subgraph rows (suminc_compress_mirror gives them; suminc_upload_sketch takes them
 * in sub_off / sub_col). */
int suminc_pr_precompute(suminc_ctx* ctx);
/* The correction at the end of the compressed phase (sum_sync_iter_worker.h:
 * 1385-1491) on the device: the exits' pending inflow is flushed first, in exit
 * passes until sum |bound| <= termcheck_threshold / 1000 (at most 256; the rest is
 * dropped, where the reference leaves all of it in bound_node_values), then every
 * non-zero spnode_datas entry is spread through its supernode's inner_value (rank)
 * and inner_delta (delta). */
int suminc_pr_correct(suminc_ctx* ctx);
/* The whole query over the uploaded sketch: the precompute (when the sketch has
 * subgraph rows), sketch rounds until sum |delta| + sum |bound| <
 * termcheck_threshold, correction, plain Ingress rounds. */
int suminc_sketch_pr_run(suminc_ctx* ctx, suminc_stats* stats);

/* The compressor with mirror vertices, from a cluster assignment (cluster_of[v]:
 * cluster id < n or SUMINC_INF_U32; clusters of min_node_num .. max_node_num members
 * taken in id order), on the host as in the reference, then the shortcut rows on
 * the device (one workgroup per supernode source, no cluster-size cap), uploaded as
 * the context's sketch (suminc_upload_sketch's layout).  Restates
 * get_init_supernode_by_clusters (mirror selection by mirror_k: an outside vertex
 * with more than mirror_k edges into (out of) a cluster becomes an in-mirror
 * (out-mirror) of it; the best of the four benefits above compress_threshold
 * accepts; compressor_base.h:3253-3611), final_build_supernode (:3616-3783),
 * build_subgraph_mirror (:3786-3856), judge_out_bound_node (:3924-4007),
 * get_nodetype_mirror (:4144-4182), sketch2csr_mirror (:4636-4899) and
 * build_iter_index_mirror (iter_compressor.h:318-491).  mirror_k = 1e8 turns mirrors
 * off (the reference's switch, compressor_base.h:3302; its default is 4,
 * examples/analytical_apps/flags.cc:62).  Uses a PageRank state: call after
 * suminc_pr_init. */
typedef struct suminc_compress_info {
  uint64_t clusters;            /* accepted clusters */
  uint64_t all_node_num;        /* vertices + mirrors */
  uint64_t supernodes;          /* supernode sources: entries and in-mirrors */
  uint64_t in_mirrors, out_mirrors;
  uint64_t absorbed;            /* unclustered mirror candidates taken into the cluster instead */
  uint64_t node_type_count[7];  /* NodeType 0..6 over the vertices */
  uint64_t is_edges, ib_edges, sync_edges, subgraph_edges;
  uint64_t inner_value_entries, inner_delta_entries;
  uint64_t shortcut_work;       /* internal edges read by the shortcut build */
  uint64_t reserved[8];
} suminc_compress_info;
int suminc_compress_mirror(suminc_ctx* ctx, const uint32_t* cluster_of, uint32_t min_node_num, uint32_t max_node_num,
                           double compress_threshold, double mirror_k, suminc_compress_info* info);
/* The last suminc_compress_mirror's products in suminc_sketch_csr's layout, into
 * caller buffers sized from its suminc_compress_info (all_node_num, is_edges, ...;
 * row pointers rows + 1 entries), plus cluster_of (all_node_num: accepted cluster
 * of every vertex and mirror, SUMINC_INF_U32 for Singles).  Any pointer may be NULL. */
typedef struct suminc_sketch_buffers {
  uint8_t* node_type;
  uint32_t* mirror_master;
  uint32_t* cluster_of;
  uint64_t* is_off;
  uint32_t* is_col;
  float* is_w;
  uint64_t* ib_off;
  uint32_t* ib_col;
  uint64_t* sync_off;
  uint32_t* sync_col;
  uint64_t* iv_off;
  uint32_t* iv_col;
  float* iv_w;
  uint64_t* id_off;
  uint32_t* id_col;
  float* id_w;
  uint64_t* sub_off;            /* all_node_num + 1 */
  uint32_t* sub_col;            /* subgraph_edges */
} suminc_sketch_buffers;
int suminc_sketch_download(suminc_ctx* ctx, suminc_sketch_buffers* buffers);
/* The cluster cache (compress_by_cluster writes it on a first run and reads it on
 * later ones, compressor_base.h:3160-3170, 3243-3249): the accepted clusters of the
 * last suminc_compress_mirror, before the final build, in the reference's layout
 * (serialize_cluster, :3081-3150; little-endian): u32 init_mirror_num, u32 cluster
 * count; per cluster five lists of u32 count + u32 ids (supernode_ids, cluster_ids,
 * supernode_source, supernode_in_mirror, supernode_out_mirror); i32 Fc[n] (-(cluster
 * + 1) for members, INT32_MAX otherwise); u32 id2spids[n] (cluster or 0xFFFFFFFF);
 * i32 -1.  The mirror sets are written in ascending order (the reference's hash
 * order is arbitrary). */
int suminc_compress_mirror_save(suminc_ctx* ctx, const char* path);
/* Reads such a file (de_serialize_cluster, :2987-3079), checks it against itself
 * (lists, Fc, id2spids, the mirror count, the end marker) and runs the final build,
 * the shortcut rows and the upload from it: the products of suminc_compress_mirror
 * without its acceptance pass. */
int suminc_compress_mirror_load(suminc_ctx* ctx, const char* path, suminc_compress_info* info);

/* ---- weighted personalized PageRank and its adjoint ------------------------------
 * No counterpart in the reference (its nearest app is the weighted personalized
 * PHP Ingress, php/php_ingress.h:15-120; nothing there has gradients).  Needs a
 * graph with SUMINC_W_F32 weights.  Forward: r = (1-d) p + d W^T r with
 * W[u,v] = w_uv / sum_k w_uk (dangling rows zero), as the delta series
 * delta_0 = (1-d) p, delta_{t+1} = d W^T delta_t, stopped when the FP64 sum|delta| <
 * termcheck_threshold.  Backward for L = sum_i c_i r_i: the same series on the
 * transpose graph from mu_0 = c, stopped when sum|mu| < termcheck_threshold * sum|c|,
 * gives lambda; dL/dp = (1-d) lambda, dL/dw_uv = (d r_u / s_u)(lambda_v - sum_k W[u,k]
 * lambda_k).  State in FP32 with Kahan sums or FP64 (pr_precision).  Vectors are n
 * floats, edge weights and their gradient m floats in out-CSR order; every pointer
 * may be host or device memory of the context's GPU.  All four synchronise.
 * suminc_wpr_backward differentiates the context's last forward (its rank and
 * weights); a caller that runs several forwards before their backwards runs the
 * forward again first (SUMINC_E_STATE when the weights or the graph changed since). */
int suminc_wpr_set_weights(suminc_ctx* ctx, const float* weights);
int suminc_wpr_forward(suminc_ctx* ctx, const float* personalization, float* rank, suminc_stats* stats);
int suminc_wpr_backward(suminc_ctx* ctx, const float* grad_rank, float* grad_personalization, float* grad_weights,
                        suminc_stats* stats);
/* Allocation-free variant of suminc_wpr_forward.  No counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that answer one personalization after another,
 * harnesses that forbid hipMalloc).  personalization and rank must be device
 * memory of the context's GPU; the call allocates nothing.  The round state lives
 * in one fixed 64 MiB device array that the library holds on every GPU it runs on
 * until the process ends: about 16 B per vertex in FP32 and 28 B per vertex in
 * FP64, so graphs up to about 4.1M (FP32) and 2.3M (FP64) vertices; larger ones
 * get SUMINC_E_TOO_LARGE.  Needs suminc_wpr_set_weights and one
 * suminc_wpr_forward in the same precision on the current weights first (they
 * prepare the row lists and per-edge coefficients), else SUMINC_E_STATE.  Every
 * call on one GPU uses that GPU's one array, which keeps the last call's contents:
 * calls that overlap on one GPU, from any context or thread, race.  The rank it
 * computes is not kept for suminc_wpr_backward (run suminc_wpr_forward before a
 * backward). */
int suminc_wpr_forward_noalloc(suminc_ctx* ctx, const float* personalization, float* rank, suminc_stats* stats);
/* FP64 copies (host memory) of the last forward's r and the last backward's lambda. */
int suminc_wpr_download(suminc_ctx* ctx, double* rank, double* adjoint);

/* ---- stats ---------------------------------------------------------------------*/
int suminc_get_stats(suminc_ctx* ctx, suminc_stats* stats);

/* ---- kernel profile -------------------------------------------------------------
 * Per-kernel device time from HIP events recorded on the context stream around
 * every launch of the named kernel that did work, plus the algorithmic bytes
 * those launches moved.  No reference counterpart (the reference times phases
 * with host timers, examples/analytical_apps/timer.h:43-80). */
#define SUMINC_KERNEL_PR_PUSH 0   /* PageRank scatter: 12 B/edge + 16 B/vertex */
#define SUMINC_KERNEL_PR_GATHER 1 /* PageRank vertex pass: 4 B/vertex + 16 B/active vertex */
#define SUMINC_KERNEL_SSSP_PUSH 2 /* SSSP relax: 16 B/edge + 16 B/vertex */
#define SUMINC_KERNEL_PR_PULL 3   /* PageRank CSC gather, plain and sketch rounds: 8 B/edge + 16 B/vertex (FP32) */
#define SUMINC_KERNEL_SSSP_SKETCH 4 /* SSSP sketch rounds: 16 B/sketch edge + 20 B/shortcut slot + 16 B/vertex */
#define SUMINC_KERNEL_BATCH_COPY 5  /* batch merge, old-edge copy: 8 B/edge per array (read + write), CSR + CSC */
#define SUMINC_KERNEL_COUNT 8
typedef struct suminc_kernel_profile {
  double launches;  /* timed launches */
  double ms;        /* summed device time */
  double alg_bytes; /* algorithmic bytes of those launches */
  double edges;     /* edges processed by those launches */
  double vertices;  /* vertices processed by those launches */
  double reserved[3];
} suminc_kernel_profile;
/* on != 0 enables event timing and clears the accumulators. */
int suminc_profile_enable(suminc_ctx* ctx, int on);
int suminc_profile_get(suminc_ctx* ctx, int kernel, suminc_kernel_profile* out);

/* ---- the reference's text formats (host only, no device work) -------------------
 * Parsed with one thread per newline-aligned chunk of a memory-mapped file.  These
 * calls take no context; on failure suminc_last_error(NULL) holds the message
 * (with the line number). */

/* Edge file: one "u v [w]" per line; empty lines and lines starting with '#' are
 * skipped, and so is a line that does not parse (counted in skipped_lines), as the
 * fragment loaders do (grape/fragment/e_fragment_loader.h:107-170; tokens as
 * grape/io/tsv_line_parser.h reads them).  weight_kind selects the third column
 * (SUMINC_W_U32: integer >= 1, SUMINC_W_F32: float, SUMINC_W_NONE: ignored).
 * directed == 0 adds both directions of every line (the loader's spec.directed).
 * Vertex ids are taken as dense ids: n = largest id + 1.  The result is a CSR with
 * sorted rows; of repeated (u, v) lines the first one is kept.  Free with
 * suminc_edges_free. */
typedef struct suminc_edges suminc_edges;
int suminc_edges_read_tsv(const char* path, int weight_kind, int directed, suminc_edges** out);
int suminc_edges_info(const suminc_edges* edges, uint32_t* n, uint64_t* m, int* weight_kind,
                      uint64_t* skipped_lines);
/* Copies the CSR out: offsets[n + 1], cols[m], weights[m] (4-byte bits); any may be NULL. */
int suminc_edges_copy(const suminc_edges* edges, uint64_t* offsets, uint32_t* cols, void* weights);
/* suminc_upload_csr of the edge list (replaces the loader + fragment build,
 * grape/fragment/loader.h:30-61 and immutable_edgecut_fragment.h:129-330). */
int suminc_edges_upload(suminc_ctx* ctx, const suminc_edges* edges);
void suminc_edges_free(suminc_edges* edges);

/* Update file: "a u v [w]" / "d u v" per line (IncFragmentBuilder's delta file,
 * grape/fragment/inc_fragment_builder.h:29-72), blank and '#' lines skipped, any
 * other op type rejected.  Adds carry the weight column when weight_kind is not
 * SUMINC_W_NONE.  directed == 0 emits both directions of every line.  *count gets
 * the number of records; they are written to ops when ops != NULL (cap must be at
 * least *count).  The records go to suminc_apply_batch unchanged. */
int suminc_updates_read_tsv(const char* path, int weight_kind, int directed, suminc_update* ops, uint64_t cap,
                            uint64_t* count);

/* Cluster file "<efile>.c_<max_node_num>": whitespace-separated records
 * "<size> <id> ... <id>", one cluster each (compressor_base.h:3153-3250).  Records
 * with fewer than min_node_num distinct ids are skipped; the kept ones are numbered
 * 0.. in file order and cluster_of[v] gets the number (SUMINC_INF_U32 for vertices
 * in none; a vertex listed in several kept clusters gets the last).  The result
 * goes to suminc_set_clusters. */
int suminc_clusters_read(const char* path, uint32_t n, uint32_t min_node_num, uint32_t* cluster_of,
                         uint32_t* n_clusters);
/* Writes cluster_of (ids < n or SUMINC_INF_U32) in the same format: one line per
 * non-empty cluster in id order, members ascending. */
int suminc_clusters_write(const char* path, uint32_t n, const uint32_t* cluster_of);
/* Community clustering for the cluster file, host only: a deterministic stand-in for
 * the reference's offline Louvain pipeline (sh/louvain_all.sh:24-49), finding
 * communities rather than the leaves of hubs.  Labels start as vertex ids; each of
 * `rounds` synchronous rounds gives every vertex with at most max_degree neighbours
 * (out- and in-edges, an edge counting once for each endpoint) the label most
 * frequent among them (a tie keeps the current label when it is among the most
 * frequent, else takes the smallest); then every label class, in vertex id order, is
 * cut into pieces of at most max_node_num members, pieces below min_node_num are
 * dropped (SUMINC_INF_U32), and the clusters are numbered by their smallest member.
 * The result does not depend on the thread count; it goes to suminc_set_clusters,
 * suminc_compress_mirror or suminc_clusters_write. */
int suminc_cluster_lpa(uint32_t n, const uint64_t* offsets, const uint32_t* cols, uint32_t max_node_num,
                       uint32_t min_node_num, uint32_t rounds, uint32_t max_degree, uint32_t* cluster_of,
                       uint32_t* n_clusters);

/* ---- binary caches of parsed inputs (host only, no device work) ------------------
 * The reference serializes a loaded fragment once under --serialization_prefix and
 * reads it back on later runs instead of parsing the text (loader.h:30-61,
 * SetSerialize).  These calls cache the edge list and the update records in a binary
 * format of this library (little-endian; an FNV-1a 64 checksum ends each file):
 *   edges    "SUMINCE1", u32 weight_kind, u32 n, u64 m, u64 skipped lines,
 *            u64 offsets[n + 1], u32 cols[m], u32 weight bits[m] (weighted graphs only)
 *   updates  "SUMINCU1", u64 count, count suminc_update records (24 bytes each)
 * A load checks the magic, the sizes against the file size, the checksum and, for an
 * edge list, the offsets (0 .. m, never decreasing) and that every column is < n. */
/* The reference's cache path "<prefix>/<efile + vfile + workers, '/' -> '_'>/frag_<fid>.s"
 * (loader.h:33-42, config.h:49) into out (cap bytes, NUL-terminated); *length gets its
 * length; out may be NULL to ask for the length. */
int suminc_serialization_path(const char* prefix, const char* efile, const char* vfile, uint32_t workers,
                              uint32_t fid, char* out, uint64_t cap, uint64_t* length);
int suminc_edges_save(const suminc_edges* edges, const char* path);
/* Free the result with suminc_edges_free. */
int suminc_edges_load(const char* path, suminc_edges** out);
int suminc_updates_save(const char* path, const suminc_update* ops, uint64_t count);
/* *count gets the record count; the records are written to ops when ops != NULL (cap
 * must be at least *count). */
int suminc_updates_load(const char* path, suminc_update* ops, uint64_t cap, uint64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* SUMINC_H_ */
