// pr.hpp -- host-side entry points of the device algorithms (called by capi.hip).
#pragma once
#include "ctx.hpp"

namespace suminc {

// graph.hip
void generate_rmat(suminc_ctx* c, int scale, int edge_factor, double a, double b, double cc, uint64_t seed,
                   int wkind, uint32_t wmin, uint32_t wmax);
void upload_csr(suminc_ctx* c, uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* col, int wkind,
                const void* w);
void download_csr(suminc_ctx* c, uint64_t* off, uint32_t* col, void* w, bool in_edges = false);

// pr.hip
void pr_alloc(suminc_ctx* c);
void pr_init(suminc_ctx* c);
void pr_run(suminc_ctx* c, suminc_stats* st);
void pr_step(suminc_ctx* c, double* residual);
void pr_download(suminc_ctx* c, double* rank, double* delta);
void pr_upload(suminc_ctx* c, const double* rank, const double* delta);
// FP64 mass balance of the current PageRank state into st->mass_total / mass_defect
void pr_mass(suminc_ctx* c, suminc_stats* st);
unsigned push_grid();
// positions of the vertices by decreasing degree (deg_off's row lengths; degree 0:
// INF32) and cols (rows rows_off, up to m_cap entries) renamed to them; scratch
// holds 16 B per vertex
void build_positions_by_degree(suminc_ctx* c, const uint32_t* deg_off, const uint32_t* rows_off,
                               const uint32_t* cols, uint64_t m_cap, DevBuf& pos, DevBuf& scratch, DevBuf& cols_c);
// whether n contributions of es bytes are large enough for compact positions
bool compact_pull(const suminc_ctx* c, uint64_t es);
// the plain pull's contribution index (nullptr: vertex ids) and CSC columns
const uint32_t* pull_pos(suminc_ctx* c);
const uint32_t* pull_cols(suminc_ctx* c);
// whether the compact positions can be carried through the coming batch (else the
// pull lists renumber them); the batch then appends the positions of the vertices
// that gained out-edges (after its CSR merge) and renames its CSC columns while
// merging them
bool pr_positions_carry(const suminc_ctx* c);
void pr_positions_append(suminc_ctx* c);
// rows with at least one entry in off (n + 1 row pointers); scratch: 8 bytes
unsigned long long nonempty_rows(suminc_ctx* c, const uint32_t* off, DevBuf& scratch);
// a binned pull layout of the adjacency (rows_off, cols) with columns < n_src, source
// windows of win_bytes / es slots (L.n_parts = 0 when it does not apply); with_eid
// keeps the row-order index of every edge in L.eid (for coefficients in L.coef)
void bin_layout_build(suminc_ctx* c, BinLayout& L, uint32_t n, const uint32_t* rows_off, const uint32_t* cols,
                      unsigned long long n_src, uint64_t es, uint64_t win_bytes, int direct_param, uint64_t m_cap,
                      bool with_eid);
// y[v] += sum over the edges (u, v) of L of x[u] (times L.coef when COEF); gate as
// the pull launches' (2: a series round, skipped once cnt is done)
template <class T, bool COEF>
void bin_layout_spmv(suminc_ctx* c, const BinLayout& L, const T* x, T* y, const Counters* cnt, int gate);
// in-degree classes of the CSC for the pull rounds: built on the stream (lengths
// stay on the device) for the current CSC, sized for up to m_cap edges
void build_pull_lists(suminc_ctx* c, uint64_t m_cap);
// the same, once per graph version
void ensure_pull_heavy(suminc_ctx* c);
// the pull launches of a PageRank round (gate 0) or reseed (gate 1), over the
// contributions at the compact positions when those are in use
template <class T>
void pull_scatter(suminc_ctx* c, int gate);

// sssp.hip
void sssp_init(suminc_ctx* c, uint32_t source);
void traversal_init(suminc_ctx* c, int app, uint32_t source);
void compress_mirror_save(suminc_ctx* c, const char* path);
void compress_mirror_load(suminc_ctx* c, const char* path, suminc_compress_info* info);
void php_run(suminc_ctx* c, uint32_t source, double damping, double* values, suminc_stats* st);
void sssp_run(suminc_ctx* c, suminc_stats* st);
void sssp_download(suminc_ctx* c, uint32_t* dist, uint32_t* parent);

// batch.hip
void generate_batch(suminc_ctx* c, uint64_t seed, uint32_t batch_id, uint64_t n_del, uint64_t n_ins, uint32_t wmin,
                    uint32_t wmax, suminc_update* ops);
void apply_batch(suminc_ctx* c, const suminc_update* ops, uint64_t n_ops, bool ops_on_device);

// weighted PageRank + adjoint (wpr.hip)
void wpr_set_weights(suminc_ctx* c, const float* w);
void wpr_forward(suminc_ctx* c, const float* p, float* r, suminc_stats* st);
void wpr_backward(suminc_ctx* c, const float* cvec, float* grad_p, float* grad_w, suminc_stats* st);
void wpr_download(suminc_ctx* c, double* r, double* lam);
// the forward over the library's fixed per-device scratch array (no allocation);
// wpr_scratch_setup takes that array's device address, once per device
void wpr_forward_noalloc(suminc_ctx* c, const float* p, float* r, suminc_stats* st);
void wpr_scratch_setup(int device);

// the reference-layout sketch (rsketch.hip)
void rsk_upload(suminc_ctx* c, const suminc_sketch_csr* s, const double* is_w64 = nullptr,
                const double* iv_w64 = nullptr, const double* id_w64 = nullptr);
void rsk_step(suminc_ctx* c, double* residual);
void rsk_correct(suminc_ctx* c);
void rsk_precompute(suminc_ctx* c);
void rsk_run(suminc_ctx* c, suminc_stats* st);
void rsk_release(suminc_ctx* c);

// compress.hip: the compressor with mirrors, its products uploaded as the sketch
void compress_mirror(suminc_ctx* c, const uint32_t* cluster_of, uint32_t min_node, uint32_t max_node,
                     double threshold, double mirror_k, suminc_compress_info* info);
void sketch_csr_download(suminc_ctx* c, suminc_sketch_buffers* b);
void mirror_release(suminc_ctx* c);

// sssp.hip: one round from the pending frontier, the next frontier's size out
void sssp_step(suminc_ctx* c, uint32_t* frontier_out);
void wpr_release(suminc_ctx* c);

}  // namespace suminc
