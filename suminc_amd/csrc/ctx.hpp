// ctx.hpp -- device context and helpers shared by the HIP translation units of
// libsuminc.so.  Written for gfx950 (CDNA4, wave64) only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/suminc.h"

namespace suminc {

// Internal failure carrying a SUMINC_E_* code; converted to a return code by
// the C ABI wrappers in capi.hip (no entry point lets an exception escape).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// A failed runtime call also leaves its error as the thread's last error, which a
// later hipGetLastError (ours or a library's) would report again: it is cleared
// before the failure is turned into a code.
#define SUMINC_HIP(expr)                                                              \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      (void)hipGetLastError();                                                        \
      throw ::suminc::Error(e_ == hipErrorOutOfMemory ? SUMINC_E_NOMEM : SUMINC_E_HIP, \
                            std::string(#expr) + ": " + hipGetErrorString(e_));       \
    }                                                                                 \
  } while (0)

// Message of the last failed call without a context on this thread (io.hip).
const char* io_last_error();
// Uploads an edge list read by suminc_edges_read_tsv (io.hip).
void upload_edges(suminc_ctx* c, const suminc_edges* e);

#define SUMINC_REQUIRE(cond, code, msg) \
  do {                                  \
    if (!(cond)) throw ::suminc::Error((code), (msg)); \
  } while (0)

// Owning device allocation; ensure() grows without preserving contents.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  void ensure(size_t bytes) {
    if (bytes <= cap && p) return;
    release();
    size_t b = bytes ? bytes : 256;
    SUMINC_HIP(hipMalloc(&p, b));
    cap = b;
  }
  // Like ensure, but a reallocation leaves 1/16 headroom: buffers whose size
  // drifts from batch to batch are not reallocated every time.
  void ensure_slack(size_t bytes) {
    if (bytes <= cap && p) return;
    ensure(bytes + bytes / 16);
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
};

// A binned pull layout (pr.hip): the in-edges of a CSR-shaped adjacency (rows =
// destinations, columns = source slots) in (source window, destination block)
// order, each packed into 32 bits ((destination - block start) << 19 | slot within
// the window); F[w * B + b] the offset of segment (w, b) (P * B + 1 values); the
// summing pass's parts (uint4: block | split << 31, 0, first and end 4-edge group of
// the block's run); the value stream (one element per edge); optionally the row
// order's index of every edge (eid) and a coefficient per edge in layout order.
struct BinLayout {
  DevBuf packed, packed2, eid, eid2, F2, F, hist, pieces, parts, vals, bounds, coef;
  uint32_t n_parts = 0, P = 0, B = 0, wl = 0, bsize = 0, direct = 0;
  uint64_t first = 0, m = 0;  // the value stream's edges: [first, m)
};

// Device-side counters block (one allocation, read back in one copy).
struct Counters {
  unsigned long long wl_count[2];   // worklist lengths (double-buffered frontiers)
  unsigned long long heavy_count;   // heavy-vertex chunk list length
  unsigned long long edges;         // edges traversed (accumulated)
  unsigned long long rounds;        // rounds executed on device
  unsigned long long done;          // stop flag set by the device
  unsigned long long aux[10];
  unsigned long long round_edges;   // active edges of the current round
  unsigned long long mode;          // PageRank scatter of the current round: 0 push, 1 pull
  unsigned long long pull_rounds;
  unsigned long long push_edges;    // active edges of push rounds
  unsigned long long push_entries;  // work-list entries of push rounds
  unsigned long long active_vertices;
  double residual;                  // last FP64 sum|delta|
  double residual_hist[2];
  double dpad[5];
  unsigned long long rc[3];         // round-indexed frontier counts (sketch SSSP rounds, reset walk)
  unsigned long long hc[3];         // round-indexed heavy-chunk counts
  unsigned long long bcopy[4];      // a batch's group counts, copied out with the counters
  unsigned long long reseed_mode;   // PageRank reseed scatter: 0 push, 1 pull
};

// Persistent scratch of the update-batch path (grows, never shrinks), so a batch
// allocates nothing in steady state.
struct BatchWs {
  DevBuf ops, keys_a, keys_b, idx_a, idx_b, rec, flags, grec;
  DevBuf g_row, g_col, g_act, g_w, g_ex, g_eoldw, g_exu, g_addu, g_expre, g_addpre, head, t_raw, t_list, t_gbeg;
  DevBuf g_pos, tg_pos;  // old position of each group's pair (CSR / CSC)
  DevBuf tstart;         // first group of every copy tile
  DevBuf tg_row, tg_col, tg_act, tg_w, tg_ex, tg_exu, tg_addu, tg_expre, tg_addpre, thead, tt_raw, tt_list, tt_gbeg;
  DevBuf tix, ttix;  // row -> touched index, all INF32 between batches
  DevBuf vts, noff_out, ncol_out, nw_out, noff_in, ncol_in, nw_in;
  DevBuf ncolc_in;  // the renamed CSC columns of the new graph (carried positions)
  DevBuf rbits, rlist;
  DevBuf val;  // validation of device-resident ops: count, first bad index
  DevBuf bc;  // device counts: [0]=G groups, [1]=touched rows (CSR), [2]=touched rows (CSC), [3]=new m
  uint32_t tix_n = 0;
  bool rbits_clear = false;  // the reset bitmap is all zero (a repair completed)
};

// Supernode sketch of the graph (mirror-free form of the reference's compressor):
// accepted clusters, node types, cluster member layout, entry (supernode) list,
// internal edges and the dense shortcut rows.  A Single's sketch row is its CSR
// row; an exit's sketch row is its CSR row without the edges inside its cluster.
struct Sketch {
  bool built = false;
  uint32_t K = 0;            // accepted clusters
  uint32_t E = 0;            // entries (one supernode each)
  uint32_t M = 0;            // clustered vertices
  uint32_t Kraw = 0;         // raw clusters in the input assignment
  uint64_t sc_len = 0;       // total length of the dense PageRank shortcut rows (one slot per member)
  uint64_t ss_len = 0;       // total length of the SSSP shortcut rows (one slot per exit)
  uint64_t ie_m = 0;         // internal edges
  uint64_t iem_n = 0;        // live entries of the internal edge list
  uint32_t max_cluster = 0;  // largest accepted cluster
  uint64_t version = ~0ull;  // graph version the structure was built for
  uint32_t RM = 0;           // vertices with a raw cluster id
  int pr_rows_for = -1;      // precision the PR shortcut rows were computed in (-1: none)
  bool ss_rows = false;      // SSSP shortcut rows exist for the current layout
  bool pr_need = true;       // some PageRank rows are invalid
  bool ss_need = true;       // some SSSP rows are invalid
  DevBuf raw_cid, raw_acc;   // input assignment (n), acceptance per raw cluster (Kraw)
  DevBuf rlist;              // vertices with a raw cluster id (RM)
  DevBuf cid, ntype, lidx;
  DevBuf base_mem, bslot;    // members in (cluster, id) order (M), base slot of each member (n); fixed
  bool batch_pending = false;  // a batch updated the internal list; merge pending
  DevBuf ext_out, ext_in;    // per clustered vertex: edges leaving / entering its cluster
  DevBuf cl_off, cl_mem, cl_nexit, cl_eoff, cl_ent, ent_of, sc_base, ss_base;
  DevBuf slot_ent;           // M: the member in this slot is an entry
  DevBuf touched;            // K: cluster touched by a batch since the last layout
  DevBuf large;              // clusters of more than 64 members (laid out one workgroup each)
  uint32_t n_large = 0;
  // internal edges: the list sorted by (base slot of dst, src), kept up to date by
  // every batch (sorted adds merged in, dead entries dropped), and the
  // per-cluster arrays the in-cluster kernels read
  DevBuf iem_key, iem_w, iem_dead, iem_key2, iem_w2, iem_pre, add_key, add_w;
  DevBuf ie_off, ie_src, ie_dst, ie_deg, ie_wt;
  DevBuf sc_R, sc_D;         // PageRank rows (reach mass / residual) per entry
  DevBuf ss_dist;            // SSSP rows (in-cluster distance to every exit) per entry
  DevBuf bound, spnode;      // PageRank sketch accumulators
  DevBuf xi_list, xi_count;  // internal edges leaving exits: (contribution slot of the exit, destination)
  DevBuf exits, exit_count;  // the exits of the current layout (sketch rounds' exit pass)
  DevBuf dead;               // K: cluster dissolved (its benefit fell to the threshold after batches)
  double threshold = 1.0;    // compress_threshold of the build (the dissolution test)
  DevBuf dirty;              // K: cluster touched since its last precompute
  DevBuf ss_dirty;           // K: SSSP precompute flags (clusters of frontier members)
  bool ss_dirty_clear = false;  // a finished query left every flag clear
  DevBuf ss_work;            // u64 x8: internal edges read by the row kernels, then the (small, large) counts of the row, precompute and correction cluster lists
  DevBuf ss_list;            // K: clusters the next per-cluster kernel visits
  DevBuf pr_valid, ss_valid; // K: shortcut rows of the cluster are current
  DevBuf sc_R_old, sc_D_old, ss_dist_old;  // rows of the previous layout
  // layout pass: per-cluster counts, their prefix (current and previous layout), type histogram
  DevBuf lay_cnt, lay_base, lay_base_old, lay_hist;
  uint32_t counts[8] = {0};  // node type histogram
};

}  // namespace suminc

struct suminc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  suminc_params prm{};
  std::string err;

  // graph: out-CSR and in-CSR (CSC), u32 row pointers, sorted by (src, dst)
  bool has_graph = false;
  uint32_t n = 0;
  uint64_t m = 0;
  int wkind = SUMINC_W_NONE;
  suminc::DevBuf out_off, out_col, out_w, in_off, in_col, in_w;

  // PageRank state
  bool pr_ready = false;
  uint64_t pr_state_id = 0;  // counts suminc_pr_init / suminc_pr_upload: a new PageRank state
  int pr_prec = SUMINC_PR_FP32_KAHAN;
  suminc::DevBuf pr_delta, pr_rank, pr_comp, pr_contrib;
  suminc::DevBuf mass_part;  // block sums of the FP64 mass balance
  // pull mode: static chunk list of high in-degree vertices, rebuilt per graph version
  suminc::DevBuf pull_heavy, pull_short, pull_wave;
  suminc::DevBuf pull_cnt;    // u64 x3: lengths of the short, wave and chunk lists
  suminc::DevBuf pull_parts;  // per-block class counts of the list build
  uint64_t pull_heavy_version = ~0ull;
  // banded pull (graphs whose contributions outgrow the caches): the wave rows'
  // in-edges split at every band boundary of the sources, and the band-major
  // chunks of the heavy rows; band_P bands of band_B source vertices
  suminc::DevBuf band_split, band_split_s, band_cnt, band_pre, band_chunks, band_off;
  uint32_t band_P = 1, band_B = 0;
  uint64_t band_W = 0, band_S = 0;
  // compact source positions (per graph version, with the pull lists): pos[u] =
  // rank of u among the vertices with out-edges, INF32 for dangling ones, and the
  // CSC columns renamed to positions, so a plain pull round gathers from an array
  // of the n_src non-dangling contributions only
  suminc::DevBuf src_pos, src_pre, in_col_c;
  bool pull_compact = false;  // the positions are in use (contributions above 32 MiB)
  // the compact positions and the renamed CSC columns agree with the current CSC
  // (pos_live); positions in use are below pos_count; a batch may carry them over
  // (pos_carried: the merge renamed its columns and appended pos_appended
  // positions for new sources) instead of renumbering every vertex
  bool pos_live = false, pos_carried = false;
  uint32_t pos_count = 0, pos_appended = 0;
  suminc::DevBuf pos_ctr;
  // tiled pull (per graph version, with the pull lists): tile table (uint4: first
  // destination, destination count | split flag << 31, CSC edge range), the CSC
  // edges of each tile grouped by source window and packed into 32 bits, and each
  // tile's window offsets (tile_P + 1 per tile)
  suminc::DevBuf tile_tab, tile_edges, tile_seg, tile_bounds, tile_edges2, tile_flat;
  uint32_t tile_n = 0, tile_P = 0, tile_win_log2 = 0, tile_bsize = 0;
  // binned pull (per graph version, with the pull lists; exclusive with the tiles)
  suminc::BinLayout bins;
  int pull_prec = -1;  // the PageRank precision the pull lists, positions and tiles were built for
  uint64_t graph_version = 0;

  // SSSP state: packed {dist:hi32, parent:lo32}
  bool sssp_ready = false;
  bool sssp_pending = false;  // initialised, not yet run
  uint32_t sssp_src = 0;
  int sssp_app = SUMINC_APP_SSSP;  // the traversal app of the SSSP state
  suminc::DevBuf sssp_dp, sssp_flag0, sssp_flag1;
  // frontier lists 0/1 and the length of list 0's seeds (set by init, appended by
  // batches, taken by the next query): owned by SSSP, so that PageRank and the
  // weighted PageRank, which reset the shared counters, leave pending seeds intact
  suminc::DevBuf sssp_wl0, sssp_wl1, sssp_seeds;

  // rounds enqueued before the first host check (last run's rounds + 1)
  int hint_reset = 6, hint_sssp = 8, hint_sssp_sketch = 8;

  // weighted PageRank / adjoint state (wpr.hip), created on first use
  void* wpr = nullptr;
  // the uploaded sketch in the reference compressor's layout (rsketch.hip), and
  // the host products of the last compression with mirrors (compress.hip)
  void* rsk = nullptr;
  void* msk = nullptr;

  // work lists and scratch
  suminc::DevBuf wl0, wl1, heavy, partial, tmp, tmp2, tmp3, tmp4, tmp5;
  suminc::BatchWs bw;
  suminc::Sketch sk;
  suminc::DevBuf counters;                  // suminc::Counters
  suminc::Counters* host_counters = nullptr;  // pinned mirror
  // pinned scratch for small device results the host reads while later kernels
  // run (copied before ev_check, read after hipEventSynchronize(ev_check))
  unsigned long long* host_scratch = nullptr;
  hipEvent_t ev_check = nullptr;

  suminc_stats stats{};

  // kernel profile (opt-in): event pairs around launches, accumulators per kernel
  bool prof_on = false;
  suminc_kernel_profile prof[SUMINC_KERNEL_COUNT]{};
  std::vector<hipEvent_t> prof_events;  // pool, reused
};

namespace suminc {

constexpr uint32_t INF32 = 0xFFFFFFFFu;
constexpr int kBlock = 256;          // threads per block for edge-parallel kernels
constexpr uint32_t kHeavyDeg = 256;  // adjacency longer than this goes to the chunk path
constexpr uint32_t kChunk = 2048;    // edges per heavy chunk
constexpr uint32_t kPullThread = 32; // pull: in-degree up to this is summed by one thread

inline unsigned grid_for(uint64_t items, int block = kBlock, unsigned cap = 1u << 30) {
  uint64_t g = (items + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

inline Counters* dev_counters(suminc_ctx* c) { return c->counters.as<Counters>(); }

// Copies the device counters block to the pinned mirror and waits.
void read_counters(suminc_ctx* c);
// A small device result read by the host while later kernels run: check_begin
// enqueues its copy into pinned memory, check_wait waits for that copy only.
void check_begin(suminc_ctx* c, const void* dev, size_t bytes);
const void* check_wait(suminc_ctx* c);
// The same for the whole counters block, into the pinned mirror.
void counters_begin(suminc_ctx* c);
void counters_wait(suminc_ctx* c);

// --- graph helpers (graph.hip) ---
void build_csc(suminc_ctx* c);  // from out-CSR
void ensure_state_sizes(suminc_ctx* c);

// --- kernel profiling (capi.hip) ---
// Event pair `slot` of the current run (grows the pool on demand).
hipEvent_t prof_event(suminc_ctx* c, size_t slot);
// Adds the elapsed time of the first `pairs` event pairs [2i, 2i+1] to kernel k.
void prof_collect(suminc_ctx* c, int k, size_t pairs, size_t first_slot = 0, size_t stride = 2);

}  // namespace suminc
