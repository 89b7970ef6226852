// cluster.hip -- community clustering for the compressor's cluster file (host only).
//
// The reference reads its clusters from "<efile>.c_<max_node_num>", written offline
// by a Louvain pipeline (sh/louvain_all.sh:24-49; read at compressor_base.h:
// 3160-3250).  This is a deterministic stand-in that finds communities rather than
// the leaves of hubs: synchronous label propagation on the undirected view of the
// graph, then every label class cut into pieces of at most max_node_num members.
//
//   labels     start as vertex ids; in each round every vertex with at most
//              max_degree neighbours takes the label most frequent among them (an
//              edge counts once for each endpoint, a reciprocal pair twice); a tie
//              keeps the current label when it is among the most frequent, else the
//              smallest; vertices with more neighbours keep their label.
//   pieces     each label class, in vertex id order, cut into consecutive pieces of
//              max_node_num; pieces below min_node_num are dropped (unclustered);
//              clusters are numbered by their smallest member.
//
// The rounds read the previous round's labels only, so the result does not depend
// on the number of threads.
#include <algorithm>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "io.hpp"

namespace suminc {
namespace {

void cluster_lpa(uint32_t n, const uint64_t* off, const uint32_t* col, uint32_t max_node, uint32_t min_node,
                 uint32_t rounds, uint32_t max_degree, uint32_t* cluster_of, uint32_t* n_clusters) {
  SUMINC_REQUIRE(off && (off[n] == 0 || col) && cluster_of, SUMINC_E_INVALID, "null input");
  SUMINC_REQUIRE(max_node >= 1 && min_node <= max_node, SUMINC_E_INVALID, "bad cluster size bounds");
  const uint64_t m = off[n];
  SUMINC_REQUIRE(off[0] == 0, SUMINC_E_INVALID, "offsets must start at 0");
  for (uint32_t v = 0; v < n; ++v) SUMINC_REQUIRE(off[v + 1] >= off[v], SUMINC_E_INVALID, "offsets decrease");
  for (uint64_t e = 0; e < m; ++e) SUMINC_REQUIRE(col[e] < n, SUMINC_E_INVALID, "column out of range");
  // the in-adjacency (transpose), sources in ascending order
  std::vector<uint64_t> ioff((size_t)n + 1, 0);
  std::vector<uint32_t> icol(m);
  for (uint64_t e = 0; e < m; ++e) ++ioff[(size_t)col[e] + 1];
  for (uint32_t v = 0; v < n; ++v) ioff[v + 1] += ioff[v];
  {
    std::vector<uint64_t> at(ioff.begin(), ioff.end() - 1);
    for (uint32_t u = 0; u < n; ++u)
      for (uint64_t e = off[u]; e < off[u + 1]; ++e) icol[at[col[e]]++] = u;
  }
  std::vector<uint32_t> label(n), next(n);
  for (uint32_t v = 0; v < n; ++v) label[v] = v;
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint32_t T = std::max(1u, std::min<uint32_t>(hw, n / 4096 + 1));
  for (uint32_t r = 0; r < rounds; ++r) {
    auto work = [&](uint32_t t) {
      std::vector<uint32_t> buf;
      const uint32_t b = (uint32_t)((uint64_t)n * t / T), e = (uint32_t)((uint64_t)n * (t + 1) / T);
      for (uint32_t v = b; v < e; ++v) {
        const uint64_t deg = (off[v + 1] - off[v]) + (ioff[v + 1] - ioff[v]);
        if (deg == 0 || deg > max_degree) {
          next[v] = label[v];
          continue;
        }
        buf.clear();
        for (uint64_t k = off[v]; k < off[v + 1]; ++k) buf.push_back(label[col[k]]);
        for (uint64_t k = ioff[v]; k < ioff[v + 1]; ++k) buf.push_back(label[icol[k]]);
        std::sort(buf.begin(), buf.end());
        uint32_t best = label[v], best_cnt = 0, own_cnt = 0;
        for (size_t i = 0; i < buf.size();) {
          size_t j = i;
          while (j < buf.size() && buf[j] == buf[i]) ++j;
          const uint32_t cnt = (uint32_t)(j - i);
          if (buf[i] == label[v]) own_cnt = cnt;
          if (cnt > best_cnt) {  // ascending labels: the first of a tie is the smallest
            best_cnt = cnt;
            best = buf[i];
          }
          i = j;
        }
        next[v] = own_cnt == best_cnt ? label[v] : best;
      }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < T; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    label.swap(next);
  }
  // label classes in id order, cut into pieces
  std::vector<uint32_t> order(n);
  for (uint32_t v = 0; v < n; ++v) order[v] = v;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return label[a] < label[b]; });
  std::vector<uint32_t> piece(n, INF32);
  std::vector<uint32_t> first;  // smallest member of every kept piece
  for (uint32_t i = 0; i < n;) {
    uint32_t j = i;
    while (j < n && label[order[j]] == label[order[i]]) ++j;
    for (uint32_t s = i; s < j; s += max_node) {
      const uint32_t e = std::min(j, s + max_node);
      if (e - s < std::max(min_node, 1u)) continue;
      const uint32_t id = (uint32_t)first.size();
      first.push_back(order[s]);  // ascending within a class
      for (uint32_t k = s; k < e; ++k) piece[order[k]] = id;
    }
    i = j;
  }
  // number the pieces by their smallest member
  std::vector<uint32_t> rank(first.size());
  for (uint32_t i = 0; i < rank.size(); ++i) rank[i] = i;
  std::sort(rank.begin(), rank.end(), [&](uint32_t a, uint32_t b) { return first[a] < first[b]; });
  std::vector<uint32_t> renum(first.size());
  for (uint32_t i = 0; i < rank.size(); ++i) renum[rank[i]] = i;
  for (uint32_t v = 0; v < n; ++v) cluster_of[v] = piece[v] == INF32 ? INF32 : renum[piece[v]];
  if (n_clusters) *n_clusters = (uint32_t)first.size();
}

}  // namespace
}  // namespace suminc

extern "C" int suminc_cluster_lpa(uint32_t n, const uint64_t* offsets, const uint32_t* cols, uint32_t max_node_num,
                                  uint32_t min_node_num, uint32_t rounds, uint32_t max_degree, uint32_t* cluster_of,
                                  uint32_t* n_clusters) {
  return suminc::host_call([&] {
    suminc::cluster_lpa(n, offsets, cols, max_node_num, min_node_num, rounds, max_degree, cluster_of, n_clusters);
  });
}
