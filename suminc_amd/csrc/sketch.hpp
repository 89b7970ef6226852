// sketch.hpp -- host entry points of the supernode sketch (sketch.hip).
#pragma once
#include "ctx.hpp"

namespace suminc {

// Internal edges of one cluster as the in-cluster kernels read them: the edges
// of cluster k (both endpoints in it) occupy [off[k], off[k + 1]).
struct Internal {
  const uint32_t* off;  // per cluster (K + 1)
  const uint32_t* src;  // local index of the source
  const uint32_t* dst;  // local index of the destination
  const uint32_t* deg;  // out-degree of the source
  const uint32_t* wt;   // integer weight (1 without weights)
};

// Per-cluster counts of the layout, and after the layout's scan their prefix: the
// first entry of cluster k and the first slot of its shortcut rows.
struct ClBase {
  unsigned long long sc;  // PageRank row slots: entries x members
  unsigned long long ss;  // SSSP row slots: entries x exits
  uint32_t ent;           // entries
  uint32_t pad;
};

inline Internal internal_of(const Sketch& s) {
  return Internal{s.ie_off.as<uint32_t>(), s.ie_src.as<uint32_t>(), s.ie_dst.as<uint32_t>(), s.ie_deg.as<uint32_t>(),
                  s.ie_wt.as<uint32_t>()};
}

void cluster_leaves(suminc_ctx* c, uint32_t max_node, uint32_t min_node, uint32_t* out);
void set_clusters(suminc_ctx* c, const uint32_t* cluster_of);
void sketch_build(suminc_ctx* c, uint32_t min_node, uint32_t max_node, double threshold);
void sketch_batch(suminc_ctx* c, const uint32_t* rows, const uint32_t* cols, const uint8_t* act, const uint8_t* ex,
                  const uint32_t* w, const unsigned* countp, uint32_t N);
void sketch_batch_done(suminc_ctx* c);
void sketch_refresh(suminc_ctx* c);
void sum_pr_run(suminc_ctx* c, suminc_stats* st);
void sum_sssp_run(suminc_ctx* c, suminc_stats* st);
void sketch_sssp_shortcut(suminc_ctx* c, uint32_t v, uint32_t cap, uint32_t* members, uint32_t* dist,
                          uint32_t* parent, uint32_t* count, uint32_t* nexit);
void sketch_info(suminc_ctx* c, suminc_sketch_info* out);
void sketch_download(suminc_ctx* c, uint32_t* cluster_of, uint8_t* node_type);
void sketch_pr_shortcut(suminc_ctx* c, uint32_t v, uint32_t cap, uint32_t* members, double* reach, double* residual,
                        uint32_t* count, uint32_t* nexit);

}  // namespace suminc
