// capi.hip -- the extern "C" boundary (include/suminc.h).  Every entry point
// validates its arguments, runs the device algorithm and converts internal
// failures into SUMINC_E_* codes with a message kept on the context.
#include <new>

#include "ctx.hpp"
#include "pr.hpp"
#include "sketch.hpp"

namespace suminc {

void read_counters(suminc_ctx* c) {
  SUMINC_HIP(hipMemcpyAsync(c->host_counters, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

// Enqueues a copy of `bytes` (at most 256) device bytes into the pinned scratch and
// records ev_check after it; check_wait() returns once the copy has landed, while
// whatever was enqueued after check_begin() keeps the device busy.
void check_begin(suminc_ctx* c, const void* dev, size_t bytes) {
  SUMINC_REQUIRE(bytes <= 256, SUMINC_E_STATE, "check copy too large");
  SUMINC_HIP(hipMemcpyAsync(c->host_scratch, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipEventRecord(c->ev_check, c->stream));
}
const void* check_wait(suminc_ctx* c) {
  SUMINC_HIP(hipEventSynchronize(c->ev_check));
  return c->host_scratch;
}
void counters_begin(suminc_ctx* c) {
  SUMINC_HIP(hipMemcpyAsync(c->host_counters, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipEventRecord(c->ev_check, c->stream));
}
void counters_wait(suminc_ctx* c) { SUMINC_HIP(hipEventSynchronize(c->ev_check)); }

hipEvent_t prof_event(suminc_ctx* c, size_t slot) {
  while (c->prof_events.size() <= slot) {
    hipEvent_t e;
    SUMINC_HIP(hipEventCreate(&e));
    c->prof_events.push_back(e);
  }
  return c->prof_events[slot];
}

void prof_collect(suminc_ctx* c, int k, size_t pairs, size_t first_slot, size_t stride) {
  if (!c->prof_on || pairs == 0) return;
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  double ms = 0;
  for (size_t i = 0; i < pairs; ++i) {
    size_t s = first_slot + i * stride;
    float t = 0;
    SUMINC_HIP(hipEventElapsedTime(&t, c->prof_events[s], c->prof_events[s + 1]));
    ms += t;
  }
  c->prof[k].launches += (double)pairs;
  c->prof[k].ms += ms;
}

template <class F>
static int guarded(suminc_ctx* c, F&& f) {
  if (!c) return SUMINC_E_INVALID;
  try {
    SUMINC_HIP(hipSetDevice(c->device));
    f();
    c->err.clear();
    return SUMINC_OK;
  } catch (const Error& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    c->err = "host allocation failed";
    return SUMINC_E_NOMEM;
  } catch (const std::exception& e) {
    c->err = e.what();
    return SUMINC_E_INVALID;
  }
}

static void check_params(const suminc_params* p) {
  SUMINC_REQUIRE(p, SUMINC_E_INVALID, "null params");
  SUMINC_REQUIRE(p->damping > 0.0 && p->damping < 1.0, SUMINC_E_INVALID, "damping must be in (0,1)");
  SUMINC_REQUIRE(p->termcheck_threshold >= 0.0, SUMINC_E_INVALID, "termcheck_threshold must be >= 0");
  SUMINC_REQUIRE(p->pr_precision == SUMINC_PR_FP64 || p->pr_precision == SUMINC_PR_FP32_KAHAN, SUMINC_E_INVALID,
                 "pr_precision must be SUMINC_PR_FP64 or SUMINC_PR_FP32_KAHAN");
}

}  // namespace suminc

using namespace suminc;

extern "C" {

void suminc_default_params(suminc_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->damping = (double)0.85f;  // PageRankIngress: const value_t d = 0.85f
  p->termcheck_threshold = 1e-6;
  p->max_rounds = 100000;
  p->pr_precision = SUMINC_PR_FP32_KAHAN;
  p->sssp_source = 0;
  p->pull_threshold_div = 3;
  p->pull_bin_direct = -1;
}

int suminc_ctx_create(int device, const suminc_params* params, suminc_ctx** out) {
  if (!out) return SUMINC_E_INVALID;
  *out = nullptr;
  suminc_ctx* c = new (std::nothrow) suminc_ctx();
  if (!c) return SUMINC_E_NOMEM;
  c->device = device;
  int rc = guarded(c, [&] {
    int count = 0;
    SUMINC_HIP(hipGetDeviceCount(&count));
    SUMINC_REQUIRE(device >= 0 && device < count, SUMINC_E_INVALID, "device index out of range");
    if (params) check_params(params);
    if (params) c->prm = *params;
    else suminc_default_params(&c->prm);
    SUMINC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    SUMINC_HIP(hipEventCreate(&c->ev_start));
    SUMINC_HIP(hipEventCreate(&c->ev_stop));
    c->counters.ensure(sizeof(Counters));
    SUMINC_HIP(hipMemsetAsync(c->counters.p, 0, sizeof(Counters), c->stream));
    void* hp = nullptr;
    SUMINC_HIP(hipHostMalloc(&hp, sizeof(Counters), hipHostMallocDefault));
    c->host_counters = reinterpret_cast<Counters*>(hp);
    std::memset(c->host_counters, 0, sizeof(Counters));
    void* hs = nullptr;
    SUMINC_HIP(hipHostMalloc(&hs, 256, hipHostMallocDefault));
    c->host_scratch = reinterpret_cast<unsigned long long*>(hs);
    std::memset(c->host_scratch, 0, 256);
    SUMINC_HIP(hipEventCreateWithFlags(&c->ev_check, hipEventDisableTiming));
    wpr_scratch_setup(device);  // device address of the allocation-free forward's scratch array
    SUMINC_HIP(hipStreamSynchronize(c->stream));
  });
  if (rc != SUMINC_OK) {
    suminc_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return SUMINC_OK;
}

int suminc_ctx_destroy(suminc_ctx* c) {
  if (!c) return SUMINC_E_INVALID;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->host_counters) (void)hipHostFree(c->host_counters);
  if (c->host_scratch) (void)hipHostFree(c->host_scratch);
  if (c->ev_check) (void)hipEventDestroy(c->ev_check);
  for (hipEvent_t e : c->prof_events) (void)hipEventDestroy(e);
  if (c->ev_start) (void)hipEventDestroy(c->ev_start);
  if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
  hipStream_t s = c->stream;
  wpr_release(c);
  rsk_release(c);
  mirror_release(c);
  delete c;  // DevBuf destructors free device memory
  if (s) (void)hipStreamDestroy(s);
  return SUMINC_OK;
}

int suminc_set_params(suminc_ctx* c, const suminc_params* p) {
  return guarded(c, [&] {
    check_params(p);
    c->prm = *p;
  });
}

const char* suminc_last_error(const suminc_ctx* c) { return c ? c->err.c_str() : suminc::io_last_error(); }

int suminc_edges_upload(suminc_ctx* c, const suminc_edges* e) {
  return guarded(c, [&] { upload_edges(c, e); });
}

int suminc_sync(suminc_ctx* c) {
  return guarded(c, [&] { SUMINC_HIP(hipStreamSynchronize(c->stream)); });
}

int suminc_get_stream(suminc_ctx* c, void** stream) {
  return guarded(c, [&] {
    SUMINC_REQUIRE(stream, SUMINC_E_INVALID, "null output");
    *stream = (void*)c->stream;
  });
}

int suminc_upload_csr(suminc_ctx* c, uint32_t n, uint64_t m, const uint64_t* offsets, const uint32_t* cols,
                      int weight_kind, const void* weights) {
  return guarded(c, [&] { upload_csr(c, n, m, offsets, cols, weight_kind, weights); });
}

int suminc_generate_rmat(suminc_ctx* c, int scale, int edge_factor, double a, double b, double cc, uint64_t seed,
                         int weight_kind, uint32_t wmin, uint32_t wmax) {
  return guarded(c, [&] { generate_rmat(c, scale, edge_factor, a, b, cc, seed, weight_kind, wmin, wmax); });
}

int suminc_graph_info(suminc_ctx* c, uint32_t* n, uint64_t* m, int* weight_kind) {
  return guarded(c, [&] {
    SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
    if (n) *n = c->n;
    if (m) *m = c->m;
    if (weight_kind) *weight_kind = c->wkind;
  });
}

int suminc_download_csr(suminc_ctx* c, uint64_t* offsets, uint32_t* cols, void* weights) {
  return guarded(c, [&] { download_csr(c, offsets, cols, weights); });
}

int suminc_download_csc(suminc_ctx* c, uint64_t* offsets, uint32_t* rows, void* weights) {
  return guarded(c, [&] { download_csr(c, offsets, rows, weights, true); });
}

int suminc_generate_batch(suminc_ctx* c, uint64_t seed, uint32_t batch_id, uint64_t n_del, uint64_t n_ins,
                          uint32_t wmin, uint32_t wmax, suminc_update* ops_out) {
  return guarded(c, [&] { generate_batch(c, seed, batch_id, n_del, n_ins, wmin, wmax, ops_out); });
}

int suminc_apply_batch(suminc_ctx* c, const suminc_update* ops, uint64_t n_ops) {
  return guarded(c, [&] { apply_batch(c, ops, n_ops, false); });
}

int suminc_apply_batch_device(suminc_ctx* c, const suminc_update* dev_ops, uint64_t n_ops) {
  return guarded(c, [&] { apply_batch(c, dev_ops, n_ops, true); });
}

int suminc_device_alloc(suminc_ctx* c, uint64_t bytes, void** dev_ptr) {
  return guarded(c, [&] {
    SUMINC_REQUIRE(dev_ptr, SUMINC_E_INVALID, "null output");
    *dev_ptr = nullptr;
    if (bytes) {
      hipError_t e = hipMalloc(dev_ptr, bytes);
      if (e != hipSuccess) {
        (void)hipGetLastError();  // not left as the thread's last error for later calls
        *dev_ptr = nullptr;
      }
      SUMINC_REQUIRE(e == hipSuccess, SUMINC_E_NOMEM, "device allocation failed");
    }
  });
}

int suminc_device_free(suminc_ctx* c, void* dev_ptr) {
  return guarded(c, [&] {
    SUMINC_HIP(hipStreamSynchronize(c->stream));
    if (dev_ptr) SUMINC_HIP(hipFree(dev_ptr));
  });
}

int suminc_copy_to_device(suminc_ctx* c, void* dev_dst, const void* host_src, uint64_t bytes) {
  return guarded(c, [&] {
    SUMINC_REQUIRE((dev_dst && host_src) || bytes == 0, SUMINC_E_INVALID, "null buffer");
    if (bytes) {
      SUMINC_HIP(hipMemcpyAsync(dev_dst, host_src, bytes, hipMemcpyDefault, c->stream));
      SUMINC_HIP(hipStreamSynchronize(c->stream));
    }
  });
}

int suminc_copy_from_device(suminc_ctx* c, void* host_dst, const void* dev_src, uint64_t bytes) {
  return guarded(c, [&] {
    SUMINC_REQUIRE((host_dst && dev_src) || bytes == 0, SUMINC_E_INVALID, "null buffer");
    if (bytes) {
      SUMINC_HIP(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, c->stream));
      SUMINC_HIP(hipStreamSynchronize(c->stream));
    }
  });
}

int suminc_pr_init(suminc_ctx* c) {
  return guarded(c, [&] { pr_init(c); });
}

int suminc_pr_run(suminc_ctx* c, suminc_stats* stats) {
  return guarded(c, [&] { pr_run(c, stats); });
}

int suminc_pr_step(suminc_ctx* c, double* residual_out) {
  return guarded(c, [&] { pr_step(c, residual_out); });
}

int suminc_pr_download(suminc_ctx* c, double* rank, double* delta) {
  return guarded(c, [&] { pr_download(c, rank, delta); });
}

int suminc_pr_upload(suminc_ctx* c, const double* rank, const double* delta) {
  return guarded(c, [&] { pr_upload(c, rank, delta); });
}

int suminc_sssp_init(suminc_ctx* c, uint32_t source) {
  return guarded(c, [&] { traversal_init(c, SUMINC_APP_SSSP, source); });
}

int suminc_traversal_init(suminc_ctx* c, int app, uint32_t source) {
  return guarded(c, [&] { traversal_init(c, app, source); });
}

int suminc_php_run(suminc_ctx* c, uint32_t source, double damping, double* values, suminc_stats* stats) {
  return guarded(c, [&] { php_run(c, source, damping, values, stats); });
}

int suminc_sssp_run(suminc_ctx* c, suminc_stats* stats) {
  return guarded(c, [&] { sssp_run(c, stats); });
}

int suminc_sssp_step(suminc_ctx* c, uint32_t* frontier_out) {
  return guarded(c, [&] { sssp_step(c, frontier_out); });
}

int suminc_upload_sketch(suminc_ctx* c, const suminc_sketch_csr* sketch) {
  return guarded(c, [&] { rsk_upload(c, sketch); });
}

int suminc_pr_step_phase(suminc_ctx* c, int phase, double* residual_out) {
  return guarded(c, [&] {
    SUMINC_REQUIRE(phase == SUMINC_PHASE_PLAIN || phase == SUMINC_PHASE_SKETCH, SUMINC_E_INVALID, "unknown phase");
    if (phase == SUMINC_PHASE_PLAIN) pr_step(c, residual_out);
    else rsk_step(c, residual_out);
  });
}

int suminc_pr_correct(suminc_ctx* c) {
  return guarded(c, [&] { rsk_correct(c); });
}

int suminc_pr_precompute(suminc_ctx* c) {
  return guarded(c, [&] { rsk_precompute(c); });
}

int suminc_sketch_pr_run(suminc_ctx* c, suminc_stats* stats) {
  return guarded(c, [&] { rsk_run(c, stats); });
}

int suminc_compress_mirror(suminc_ctx* c, const uint32_t* cluster_of, uint32_t min_node_num, uint32_t max_node_num,
                           double compress_threshold, double mirror_k, suminc_compress_info* info) {
  return guarded(c, [&] {
    compress_mirror(c, cluster_of, min_node_num, max_node_num, compress_threshold, mirror_k, info);
  });
}

int suminc_compress_mirror_save(suminc_ctx* c, const char* path) {
  return guarded(c, [&] { compress_mirror_save(c, path); });
}

int suminc_compress_mirror_load(suminc_ctx* c, const char* path, suminc_compress_info* info) {
  return guarded(c, [&] { compress_mirror_load(c, path, info); });
}

int suminc_sketch_download(suminc_ctx* c, suminc_sketch_buffers* buffers) {
  return guarded(c, [&] { sketch_csr_download(c, buffers); });
}

int suminc_sssp_download(suminc_ctx* c, uint32_t* dist, uint32_t* parent) {
  return guarded(c, [&] { sssp_download(c, dist, parent); });
}

int suminc_cluster_leaves(suminc_ctx* c, uint32_t max_node_num, uint32_t min_node_num, uint32_t* cluster_of_out) {
  return guarded(c, [&] { cluster_leaves(c, max_node_num, min_node_num, cluster_of_out); });
}

int suminc_set_clusters(suminc_ctx* c, const uint32_t* cluster_of) {
  return guarded(c, [&] { set_clusters(c, cluster_of); });
}

int suminc_sum_build(suminc_ctx* c, uint32_t min_node_num, uint32_t max_node_num, double compress_threshold) {
  return guarded(c, [&] {
    SUMINC_REQUIRE(max_node_num >= 1 && min_node_num <= max_node_num, SUMINC_E_INVALID, "bad cluster size bounds");
    c->sk.raw_acc.release();  // a fresh build decides acceptance again
    c->sk.built = false;
    sketch_build(c, min_node_num, max_node_num, compress_threshold);
  });
}

int suminc_sum_info(suminc_ctx* c, suminc_sketch_info* info) {
  return guarded(c, [&] { sketch_info(c, info); });
}

int suminc_sum_download(suminc_ctx* c, uint32_t* cluster_of, uint8_t* node_type) {
  return guarded(c, [&] { sketch_download(c, cluster_of, node_type); });
}

int suminc_sum_pr_shortcut(suminc_ctx* c, uint32_t entry, uint32_t cap, uint32_t* members, double* reach,
                           double* residual, uint32_t* count, uint32_t* nexit) {
  return guarded(c, [&] { sketch_pr_shortcut(c, entry, cap, members, reach, residual, count, nexit); });
}

int suminc_sum_pr_run(suminc_ctx* c, suminc_stats* stats) {
  return guarded(c, [&] { sum_pr_run(c, stats); });
}

int suminc_sum_sssp_shortcut(suminc_ctx* c, uint32_t entry, uint32_t cap, uint32_t* members, uint32_t* dist,
                             uint32_t* parent, uint32_t* count, uint32_t* nexit) {
  return guarded(c, [&] { sketch_sssp_shortcut(c, entry, cap, members, dist, parent, count, nexit); });
}

int suminc_sum_sssp_run(suminc_ctx* c, suminc_stats* stats) {
  return guarded(c, [&] { sum_sssp_run(c, stats); });
}

int suminc_wpr_set_weights(suminc_ctx* c, const float* weights) {
  return guarded(c, [&] { wpr_set_weights(c, weights); });
}

int suminc_wpr_forward(suminc_ctx* c, const float* personalization, float* rank, suminc_stats* stats) {
  return guarded(c, [&] { wpr_forward(c, personalization, rank, stats); });
}

int suminc_wpr_backward(suminc_ctx* c, const float* grad_rank, float* grad_personalization, float* grad_weights,
                        suminc_stats* stats) {
  return guarded(c, [&] { wpr_backward(c, grad_rank, grad_personalization, grad_weights, stats); });
}

int suminc_wpr_forward_noalloc(suminc_ctx* c, const float* personalization, float* rank, suminc_stats* stats) {
  return guarded(c, [&] { wpr_forward_noalloc(c, personalization, rank, stats); });
}

int suminc_wpr_download(suminc_ctx* c, double* rank, double* adjoint) {
  return guarded(c, [&] { wpr_download(c, rank, adjoint); });
}

int suminc_get_stats(suminc_ctx* c, suminc_stats* stats) {
  return guarded(c, [&] {
    SUMINC_REQUIRE(stats, SUMINC_E_INVALID, "null output");
    *stats = c->stats;
  });
}

int suminc_profile_enable(suminc_ctx* c, int on) {
  return guarded(c, [&] {
    c->prof_on = on != 0;
    for (auto& p : c->prof) p = suminc_kernel_profile{};
  });
}

int suminc_profile_get(suminc_ctx* c, int kernel, suminc_kernel_profile* out) {
  return guarded(c, [&] {
    SUMINC_REQUIRE(out, SUMINC_E_INVALID, "null output");
    SUMINC_REQUIRE(kernel >= 0 && kernel < SUMINC_KERNEL_COUNT, SUMINC_E_INVALID, "kernel id out of range");
    *out = c->prof[kernel];
  });
}

}  // extern "C"
