// sketch.hpp -- host entry points of the supernode sketch (sketch.hip).
#pragma once
#include "ctx.hpp"

namespace suminc {

void cluster_leaves(suminc_ctx* c, uint32_t max_node, uint32_t min_node, uint32_t* out);
void set_clusters(suminc_ctx* c, const uint32_t* cluster_of);
void sketch_build(suminc_ctx* c, uint32_t min_node, uint32_t max_node, double threshold);
void sketch_mark_dirty(suminc_ctx* c, const uint32_t* rows, const uint32_t* cols, const unsigned* countp, uint32_t N);
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
