/* worker_phases.c -- drives libsuminc through the reference worker's phase order
 * with nothing but the C ABI, the way SumSyncIterWorker::Query would after the
 * binding: compressor products in (suminc_compress_mirror on one context, its
 * download uploaded into a second through suminc_upload_sketch), then
 *   precompute (sum_sync_iter_worker.h:387-435)
 *   compressed rounds until the residual falls below T (:905-1142, 1372)
 *   correction (:1385-1491)
 *   Ingress rounds until the residual falls below T (:840-904)
 * and, after an update batch (reloadGraph + AmendValue, :1511-1694), the same again
 * over the compressor's new products; SSSP single steps until the frontier is
 * empty (sum_sync_traversal_worker.h:1136-1142).  Checks against the library's
 * plain PageRank and SSSP on a third context.  Prints one line per check and
 * "worker_phases: PASS" at the end; exit status 0 only then.
 *
 * Build: gcc -std=c11 -O2 -I include tests/c/worker_phases.c -L suminc_amd -lsuminc
 *        -Wl,-rpath,<abs path of suminc_amd> -lm */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "suminc.h"

#define CHECK(call)                                                                 \
  do {                                                                              \
    int rc_ = (call);                                                               \
    if (rc_ != SUMINC_OK) {                                                         \
      fprintf(stderr, "%s:%d: %s -> %d\n", __FILE__, __LINE__, #call, rc_);         \
      return 1;                                                                     \
    }                                                                               \
  } while (0)
#define EXPECT(cond, ...)                                                           \
  do {                                                                              \
    if (!(cond)) {                                                                  \
      fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__);                        \
      fprintf(stderr, __VA_ARGS__);                                                 \
      fprintf(stderr, "\n");                                                        \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

/* planted clusters that want mirrors: 24 clusters of 32 members, 8 out-edges each
 * inside, per cluster 3 members with 8 edges into another cluster and 8 members of
 * that cluster with an edge to one member, 64 unclustered vertices, noise edges */
enum { NCL = 24, SZ = 32, NFREE = 64, FAN = 8 };

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd(uint32_t bound) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (uint32_t)(rng_state % bound);
}

typedef struct { uint32_t u, v; } edge;

static int cmp_edge(const void* a, const void* b) {
  const edge* x = (const edge*)a;
  const edge* y = (const edge*)b;
  if (x->u != y->u) return x->u < y->u ? -1 : 1;
  if (x->v != y->v) return x->v < y->v ? -1 : 1;
  return 0;
}

static uint32_t build_graph(edge** out_e, uint64_t* out_m) {
  const uint32_t n = NCL * SZ + NFREE;
  size_t cap = 1 << 16, m = 0;
  edge* e = (edge*)malloc(cap * sizeof(edge));
#define ADD(a, b)                                      \
  do {                                                 \
    if ((a) != (b)) {                                  \
      if (m == cap) e = (edge*)realloc(e, (cap *= 2) * sizeof(edge)); \
      e[m].u = (a);                                    \
      e[m].v = (b);                                    \
      ++m;                                             \
    }                                                  \
  } while (0)
  for (uint32_t c = 0; c < NCL; ++c)
    for (uint32_t i = 0; i < SZ; ++i)
      for (int k = 0; k < FAN; ++k) ADD(c * SZ + i, c * SZ + rnd(SZ));
  for (uint32_t c = 0; c < NCL; ++c)
    for (int l = 0; l < 3; ++l) {
      const uint32_t y = rnd(NCL);
      if (y == c) continue;
      const uint32_t w = c * SZ + rnd(SZ), t = c * SZ + rnd(SZ);
      for (int k = 0; k < FAN; ++k) ADD(w, y * SZ + rnd(SZ));
      for (int k = 0; k < FAN; ++k) ADD(y * SZ + rnd(SZ), t);
    }
  for (uint32_t f = 0; f < NFREE; ++f) {
    const uint32_t u = NCL * SZ + f, y = rnd(NCL);
    for (int k = 0; k < (f % 2 ? 2 : FAN); ++k) {
      if (f % 4 < 2) ADD(u, y * SZ + rnd(SZ));
      else ADD(y * SZ + rnd(SZ), u);
    }
    ADD(u, rnd(n));
    ADD(rnd(n), u);
  }
  for (uint32_t u = 0; u < NCL * SZ; ++u)
    if (rnd(10) == 0) ADD(u, rnd(n));
#undef ADD
  qsort(e, m, sizeof(edge), cmp_edge);
  size_t k = 0;
  for (size_t i = 0; i < m; ++i)
    if (k == 0 || cmp_edge(&e[i], &e[k - 1]) != 0) e[k++] = e[i];
  *out_e = e;
  *out_m = k;
  return n;
}

static int upload(suminc_ctx* ctx, uint32_t n, const edge* e, uint64_t m) {
  uint64_t* off = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
  uint32_t* col = (uint32_t*)malloc(m * sizeof(uint32_t));
  for (uint64_t i = 0; i < m; ++i) {
    off[e[i].u + 1]++;
    col[i] = e[i].v;
  }
  for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
  const int rc = suminc_upload_csr(ctx, n, m, off, col, SUMINC_W_NONE, NULL);
  free(off);
  free(col);
  return rc;
}

static double l1(const double* a, const double* b, uint32_t n) {
  double s = 0;
  for (uint32_t i = 0; i < n; ++i) s += fabs(a[i] - b[i]);
  return s;
}

/* the worker's compressed query, phase by phase; returns the rounds of each phase */
static int query_by_phases(suminc_ctx* ctx, double T, int* sketch_rounds, int* plain_rounds) {
  double res = 0;
  *sketch_rounds = *plain_rounds = 0;
  CHECK(suminc_pr_precompute(ctx));
  for (;;) {
    CHECK(suminc_pr_step_phase(ctx, SUMINC_PHASE_SKETCH, &res));
    if (res < T) break;
    if (++*sketch_rounds > 100000) return 1;
  }
  CHECK(suminc_pr_correct(ctx));
  for (;;) {
    CHECK(suminc_pr_step_phase(ctx, SUMINC_PHASE_PLAIN, &res));
    if (res < T) break;
    if (++*plain_rounds > 100000) return 1;
  }
  return 0;
}

typedef struct {
  suminc_sketch_buffers b;
  suminc_sketch_csr s;
} sketch_copy;

static void* zalloc(uint64_t count, size_t size) { return calloc(count ? count : 1, size); }

/* suminc_sketch_download into fresh buffers, pointed at by a suminc_sketch_csr */
static int copy_sketch(suminc_ctx* ctx, uint32_t n, const suminc_compress_info* ci, sketch_copy* k) {
  const uint64_t a = ci->all_node_num;
  suminc_sketch_buffers* b = &k->b;
  memset(k, 0, sizeof(*k));
  b->node_type = (uint8_t*)zalloc(n, 1);
  b->mirror_master = (uint32_t*)zalloc(a - n, 4);
  b->cluster_of = (uint32_t*)zalloc(a, 4);
  b->is_off = (uint64_t*)zalloc(a + 1, 8);
  b->is_col = (uint32_t*)zalloc(ci->is_edges, 4);
  b->is_w = (float*)zalloc(ci->is_edges, 4);
  b->ib_off = (uint64_t*)zalloc(n + 1, 8);
  b->ib_col = (uint32_t*)zalloc(ci->ib_edges, 4);
  b->sync_off = (uint64_t*)zalloc(a + 1, 8);
  b->sync_col = (uint32_t*)zalloc(ci->sync_edges, 4);
  b->iv_off = (uint64_t*)zalloc(a + 1, 8);
  b->iv_col = (uint32_t*)zalloc(ci->inner_value_entries, 4);
  b->iv_w = (float*)zalloc(ci->inner_value_entries, 4);
  b->id_off = (uint64_t*)zalloc(a + 1, 8);
  b->id_col = (uint32_t*)zalloc(ci->inner_delta_entries, 4);
  b->id_w = (float*)zalloc(ci->inner_delta_entries, 4);
  b->sub_off = (uint64_t*)zalloc(a + 1, 8);
  b->sub_col = (uint32_t*)zalloc(ci->subgraph_edges, 4);
  CHECK(suminc_sketch_download(ctx, b));
  suminc_sketch_csr* s = &k->s;
  s->n = n;
  s->all_node_num = (uint32_t)a;
  s->node_type = b->node_type;
  s->mirror_master = b->mirror_master;
  s->is_off = b->is_off; s->is_col = b->is_col; s->is_w = b->is_w;
  s->ib_off = b->ib_off; s->ib_col = b->ib_col;
  s->sync_off = b->sync_off; s->sync_col = b->sync_col;
  s->iv_off = b->iv_off; s->iv_col = b->iv_col; s->iv_w = b->iv_w;
  s->id_off = b->id_off; s->id_col = b->id_col; s->id_w = b->id_w;
  s->sub_off = b->sub_off; s->sub_col = b->sub_col;
  return 0;
}

static void free_sketch(sketch_copy* k) {
  void* p[] = {k->b.node_type, k->b.mirror_master, k->b.cluster_of, k->b.is_off, k->b.is_col, k->b.is_w,
               k->b.ib_off, k->b.ib_col, k->b.sync_off, k->b.sync_col, k->b.iv_off, k->b.iv_col, k->b.iv_w,
               k->b.id_off, k->b.id_col, k->b.id_w, k->b.sub_off, k->b.sub_col};
  for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); ++i) free(p[i]);
}

int main(void) {
  edge* e = NULL;
  uint64_t m = 0;
  const uint32_t n = build_graph(&e, &m);
  suminc_params prm;
  suminc_default_params(&prm);
  prm.pr_precision = SUMINC_PR_FP64;
  prm.termcheck_threshold = 1e-10 * n;
  const double T = prm.termcheck_threshold, d = prm.damping;
  suminc_ctx *A = NULL, *B = NULL, *P = NULL;
  CHECK(suminc_ctx_create(0, &prm, &A));
  CHECK(suminc_ctx_create(0, &prm, &B));
  CHECK(suminc_ctx_create(0, &prm, &P));
  CHECK(upload(A, n, e, m));
  CHECK(upload(B, n, e, m));
  CHECK(upload(P, n, e, m));
  uint32_t* cl = (uint32_t*)malloc(n * sizeof(uint32_t));
  for (uint32_t v = 0; v < n; ++v) cl[v] = v < NCL * SZ ? v / SZ : SUMINC_INF_U32;

  /* the compressor's products (mirror_k 4, the reference's default) */
  suminc_compress_info ci;
  CHECK(suminc_pr_init(A));
  CHECK(suminc_compress_mirror(A, cl, 2, 1024, 1.0, 4.0, &ci));
  printf("compress: n %u m %llu clusters %llu all_node_num %llu in-mirrors %llu out-mirrors %llu types",
         n, (unsigned long long)m, (unsigned long long)ci.clusters, (unsigned long long)ci.all_node_num,
         (unsigned long long)ci.in_mirrors, (unsigned long long)ci.out_mirrors);
  for (int t = 0; t < 7; ++t) printf(" %llu", (unsigned long long)ci.node_type_count[t]);
  printf("\n");
  EXPECT(ci.clusters > 0 && ci.in_mirrors > 0 && ci.out_mirrors > 0, "no mirrors on the planted graph");
  EXPECT(ci.node_type_count[5] > 0 && ci.node_type_count[6] > 0, "no master types");

  /* reference: the plain Ingress PageRank */
  double *ra = (double*)malloc(n * 8), *rb = (double*)malloc(n * 8), *rp = (double*)malloc(n * 8);
  double* dl = (double*)malloc(n * 8);
  suminc_stats st;
  CHECK(suminc_pr_init(P));
  CHECK(suminc_pr_run(P, &st));
  CHECK(suminc_pr_download(P, rp, dl));

  /* A: the phase order by hand */
  int sr = 0, pr = 0;
  CHECK(suminc_pr_init(A));
  if (query_by_phases(A, T, &sr, &pr)) return 1;
  CHECK(suminc_pr_download(A, ra, dl));
  double err = l1(ra, rp, n);
  printf("phases: sketch rounds %d, plain rounds %d, L1 vs plain %.3e\n", sr, pr, err);
  EXPECT(sr > 0 && err < 4 * T / (1 - d), "phase-driven query off the plain fixed point");

  /* B: the downloaded products through suminc_upload_sketch, whole query */
  sketch_copy k;
  if (copy_sketch(A, n, &ci, &k)) return 1;
  CHECK(suminc_upload_sketch(B, &k.s));
  CHECK(suminc_pr_init(B));
  CHECK(suminc_sketch_pr_run(B, &st));
  CHECK(suminc_pr_download(B, rb, dl));
  err = l1(rb, rp, n);
  printf("uploaded sketch: sketch rounds %.0f, residual rounds %.0f, mass defect %.2e, L1 vs plain %.3e\n",
         st.sketch_rounds, st.residual_rounds, st.mass_defect, err);
  /* FP32 weights in the download: 2^-24 relative per use, within (1-d)^-2 of mass */
  EXPECT(err < 4 * T / (1 - d) + 8 * ldexp(1.0, -24) / ((1 - d) * (1 - d)), "uploaded sketch off the fixed point");
  free_sketch(&k);

  /* an update batch: the stale sketch is refused, the compressor runs again */
  enum { NOPS = 200 };
  suminc_update ops[NOPS];
  CHECK(suminc_generate_batch(A, 7, 0, NOPS / 2, NOPS / 2, 1, 1, ops));
  CHECK(suminc_apply_batch(A, ops, NOPS));
  CHECK(suminc_apply_batch(P, ops, NOPS));
  double res = 0;
  EXPECT(suminc_pr_step_phase(A, SUMINC_PHASE_SKETCH, &res) == SUMINC_E_STATE, "stale sketch accepted");
  CHECK(suminc_compress_mirror(A, cl, 2, 1024, 1.0, 4.0, &ci));
  if (query_by_phases(A, T, &sr, &pr)) return 1;  /* from the reseeded state */
  CHECK(suminc_pr_download(A, ra, dl));
  CHECK(suminc_pr_init(P));
  CHECK(suminc_pr_run(P, &st));
  CHECK(suminc_pr_download(P, rp, dl));
  err = l1(ra, rp, n);
  printf("after batch: sketch rounds %d, plain rounds %d, L1 vs a plain run from scratch %.3e\n", sr, pr, err);
  EXPECT(err < 4 * T / (1 - d), "incremental phase-driven query off the new fixed point");

  /* SSSP one step at a time vs the whole run */
  uint32_t *da = (uint32_t*)malloc(n * 4), *pa = (uint32_t*)malloc(n * 4);
  uint32_t *dp = (uint32_t*)malloc(n * 4), *pp = (uint32_t*)malloc(n * 4);
  CHECK(suminc_sssp_init(A, 0));
  uint32_t frontier = 1;
  int steps = 0;
  while (frontier) {
    CHECK(suminc_sssp_step(A, &frontier));
    EXPECT(++steps < 100000, "SSSP steps do not end");
  }
  CHECK(suminc_sssp_download(A, da, pa));
  CHECK(suminc_sssp_init(P, 0));
  CHECK(suminc_sssp_run(P, &st));
  CHECK(suminc_sssp_download(P, dp, pp));
  EXPECT(memcmp(da, dp, n * 4) == 0, "SSSP distances differ");
  uint32_t reached = 0;
  for (uint32_t v = 0; v < n; ++v) reached += da[v] != SUMINC_INF_U32;
  printf("sssp: %d steps, %u reached, equal to the whole run\n", steps, reached);

  CHECK(suminc_ctx_destroy(A));
  CHECK(suminc_ctx_destroy(B));
  CHECK(suminc_ctx_destroy(P));
  free(e); free(cl); free(ra); free(rb); free(rp); free(dl); free(da); free(pa); free(dp); free(pp);
  printf("worker_phases: PASS\n");
  return 0;
}
