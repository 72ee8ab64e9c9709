// C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY; see
// gas_oracle.h).  Each driver restates the reference app's run<Engine> loop:
// vertex initialisation, setGraph, setActive, the GAS loop, and returns the
// loop-trip count.  `stats` (may be NULL) receives
// [iterations, sum #A, sum E_in(A), sum #C, sum E_out(C), loop_ns].
#include <stdint.h>

#include <chrono>
#include <climits>
#include <cstring>
#include <vector>

#include "gas_oracle.h"

using namespace oracle;

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

template <class E>
void put_stats(const E& eng, int64_t ns, int64_t* stats) {
  if (!stats) return;
  stats[0] = eng.stats.iterations;
  stats[1] = eng.stats.active;
  stats[2] = eng.stats.in_edges;
  stats[3] = eng.stats.changed;
  stats[4] = eng.stats.out_edges;
  stats[5] = ns;
}

// bfs.cu:93-134: level counter starts at 0 with {source} active; gather and
// scatter flags false; one trip per level.
template <class E>
int32_t bfs_loop(E& eng, int32_t source) {
  eng.setActive(source, source + 1);
  int32_t iteration = 0;
  while (eng.countActive()) {
    eng.gather(false);
    eng.apply(iteration);
    eng.scatterActivate(false);
    eng.nextIter();
    ++iteration;
  }
  return iteration;
}

// sssp.cu:69-111: all vertices active at the start.
template <class E>
int32_t all_active_loop(E& eng, int32_t nV) {
  eng.setActive(0, nV);
  int32_t it = 0;
  while (eng.countActive()) {
    eng.gather();
    eng.apply(it);
    eng.scatterActivate();
    eng.nextIter();
    ++it;
  }
  return it;
}

}  // namespace

extern "C" {

int oracle_bfs(int32_t nV, int64_t nE, const int32_t* srcs, const int32_t* dsts, int32_t source,
               int32_t* depth, int64_t* stats) {
  std::vector<BFS::VertexData> vd((size_t)nV);
  for (auto& v : vd) v.depth = -1;
  Engine<BFS> eng;
  eng.setGraph(nV, vd.data(), nE, nullptr, srcs, dsts);
  int64_t t0 = now_ns();
  int32_t it = bfs_loop(eng, source);
  put_stats(eng, now_ns() - t0, stats);
  for (int32_t i = 0; i < nV; ++i) depth[i] = vd[i].depth;
  return it;
}

int oracle_bfs_csr(int32_t nV, const int32_t* off_out, const int32_t* dsts, const int32_t* off_in,
                   const int32_t* srcs, int32_t source, int32_t* depth, int64_t* stats) {
  std::vector<BFS::VertexData> vd((size_t)nV);
  for (auto& v : vd) v.depth = -1;
  Engine<BFS> eng;
  eng.setGraphCSR(nV, vd.data(), off_out, dsts, off_in, srcs, nullptr);
  int64_t t0 = now_ns();
  int32_t it = bfs_loop(eng, source);
  put_stats(eng, now_ns() - t0, stats);
  for (int32_t i = 0; i < nV; ++i) depth[i] = vd[i].depth;
  return it;
}

// weights in edge-list order
int oracle_sssp(int32_t nV, int64_t nE, const int32_t* srcs, const int32_t* dsts,
                const int32_t* weights, int32_t source, int32_t* dist, int64_t* stats) {
  std::vector<int> vd((size_t)nV, SSSP::gatherZero);
  vd[source] = 0;
  std::vector<int> ed(weights, weights + nE);
  Engine<SSSP> eng;
  eng.setGraph(nV, vd.data(), nE, ed.data(), srcs, dsts);
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  std::memcpy(dist, vd.data(), sizeof(int32_t) * nV);
  return it;
}

// weights: in, edge-list order; out: the weights after the run (scatter)
int oracle_sssp_congestion(int32_t nV, int64_t nE, const int32_t* srcs, const int32_t* dsts,
                           int32_t* weights, int32_t source, int32_t* dist, int64_t* stats) {
  std::vector<int> vd((size_t)nV, SSSPCongestion::gatherZero);
  vd[source] = 0;
  Engine<SSSPCongestion> eng;
  eng.setGraph(nV, vd.data(), nE, weights, srcs, dsts);
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  std::memcpy(dist, vd.data(), sizeof(int32_t) * nV);
  return it;
}

// weights_csc in CSC order
int oracle_sssp_csr(int32_t nV, const int32_t* off_out, const int32_t* dsts, const int32_t* off_in,
                    const int32_t* srcs, const int32_t* weights_csc, int32_t source, int32_t* dist,
                    int64_t* stats) {
  std::vector<int> vd((size_t)nV, SSSP::gatherZero);
  vd[source] = 0;
  std::vector<int> ed(weights_csc, weights_csc + off_in[nV]);
  Engine<SSSP> eng;
  eng.setGraphCSR(nV, vd.data(), off_out, dsts, off_in, srcs, ed.data());
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  std::memcpy(dist, vd.data(), sizeof(int32_t) * nV);
  return it;
}

// Widest paths from `source` (INT_MAX there, 0 elsewhere), all vertices
// active first, as sssp.cu's loop; capacities in edge-list order.
int oracle_widest(int32_t nV, int64_t nE, const int32_t* srcs, const int32_t* dsts,
                  const int32_t* caps, int32_t source, int32_t* width, int64_t* stats) {
  std::vector<int> vd((size_t)nV, WidestPath::gatherZero);
  vd[source] = INT_MAX;
  std::vector<int> ed(caps, caps + nE);
  Engine<WidestPath> eng;
  eng.setGraph(nV, vd.data(), nE, ed.data(), srcs, dsts);
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  std::memcpy(width, vd.data(), sizeof(int32_t) * nV);
  return it;
}

int oracle_cc(int32_t nV, int64_t nE, const int32_t* srcs, const int32_t* dsts, int32_t* label,
              int64_t* stats) {
  std::vector<int> vd((size_t)nV);
  for (int32_t i = 0; i < nV; ++i) vd[i] = i;
  Engine<CC> eng;
  eng.setGraph(nV, vd.data(), nE, nullptr, srcs, dsts);
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  std::memcpy(label, vd.data(), sizeof(int32_t) * nV);
  return it;
}

int oracle_cc_csr(int32_t nV, const int32_t* off_out, const int32_t* dsts, const int32_t* off_in,
                  const int32_t* srcs, int32_t* label, int64_t* stats) {
  std::vector<int> vd((size_t)nV);
  for (int32_t i = 0; i < nV; ++i) vd[i] = i;
  Engine<CC> eng;
  eng.setGraphCSR(nV, vd.data(), off_out, dsts, off_in, srcs, nullptr);
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  std::memcpy(label, vd.data(), sizeof(int32_t) * nV);
  return it;
}

// pagerank.cu:85-101: ranks start at pageConst, numOutEdges from the CSR.
int oracle_pagerank(int32_t nV, int64_t nE, const int32_t* srcs, const int32_t* dsts, float* rank,
                    int64_t* stats) {
  std::vector<PageRank::VertexData> vd((size_t)nV);
  Engine<PageRank> eng;
  eng.setGraph(nV, vd.data(), nE, nullptr, srcs, dsts);
  const std::vector<int32_t>& off = eng.outOffsets();
  for (int32_t i = 0; i < nV; ++i) {
    vd[i].rank = PageRank::pageConst;
    vd[i].numOutEdges = off[i + 1] - off[i];
  }
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  for (int32_t i = 0; i < nV; ++i) rank[i] = vd[i].rank;
  return it;
}

// The same loop in FP64 (PageRank64): ranks start at (double)0.15f.
int oracle_pagerank64(int32_t nV, int64_t nE, const int32_t* srcs, const int32_t* dsts,
                      double* rank, int64_t* stats) {
  std::vector<PageRank64::VertexData> vd((size_t)nV);
  Engine<PageRank64> eng;
  eng.setGraph(nV, vd.data(), nE, nullptr, srcs, dsts);
  const std::vector<int32_t>& off = eng.outOffsets();
  for (int32_t i = 0; i < nV; ++i) {
    vd[i].rank = PageRank64::pageConst;
    vd[i].numOutEdges = off[i + 1] - off[i];
  }
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  for (int32_t i = 0; i < nV; ++i) rank[i] = vd[i].rank;
  return it;
}

int oracle_pagerank_csr(int32_t nV, const int32_t* off_out, const int32_t* dsts,
                        const int32_t* off_in, const int32_t* srcs, float* rank, int64_t* stats) {
  std::vector<PageRank::VertexData> vd((size_t)nV);
  for (int32_t i = 0; i < nV; ++i) {
    vd[i].rank = PageRank::pageConst;
    vd[i].numOutEdges = off_out[i + 1] - off_out[i];
  }
  Engine<PageRank> eng;
  eng.setGraphCSR(nV, vd.data(), off_out, dsts, off_in, srcs, nullptr);
  int64_t t0 = now_ns();
  int32_t it = all_active_loop(eng, nV);
  put_stats(eng, now_ns() - t0, stats);
  for (int32_t i = 0; i < nV; ++i) rank[i] = vd[i].rank;
  return it;
}

// All-active PageRank recurrence of pagerank.cu:46-66 in FP64:
// r'[v] = 0.15 + 0.85 * sum_{u->v} r[u] / outdeg[u], r starts at 0.15, for
// `iterations` steps (the precision form's reference).  l1[k] (may be NULL)
// receives sum_v |r_k+1[v] - r_k[v]| per step.  CSC input.
int oracle_pagerank_fp64_csr(int32_t nV, const int32_t* off_out, const int32_t* off_in,
                             const int32_t* srcs, int32_t iterations, double* rank, double* l1) {
  std::vector<double> r((size_t)nV, 0.15), c((size_t)nV), nr((size_t)nV);
  for (int32_t k = 0; k < iterations; ++k) {
    for (int32_t u = 0; u < nV; ++u) {
      const int32_t d = off_out[u + 1] - off_out[u];
      c[u] = d ? r[u] / (double)d : 0.0;
    }
    double diff = 0;
    for (int32_t v = 0; v < nV; ++v) {
      double s = 0;
      for (int32_t p = off_in[v]; p < off_in[v + 1]; ++p) s += c[srcs[p]];
      nr[v] = 0.15 + 0.85 * s;
      diff += std::fabs(nr[v] - r[v]);
    }
    r.swap(nr);
    if (l1) l1[k] = diff;
  }
  std::memcpy(rank, r.data(), sizeof(double) * nV);
  return iterations;
}

// CSR/CSC of an edge list with the oracle's builder (util.cuh:79-121).
void oracle_build_csr(int32_t nV, int64_t nE, const int32_t* rows, const int32_t* cols,
                      int32_t* offsets, int32_t* adj) {
  std::vector<int32_t> off, a, perm;
  edge_list_to_csr(nV, nE, rows, cols, off, a, perm);
  std::memcpy(offsets, off.data(), sizeof(int32_t) * (nV + 1));
  std::memcpy(adj, a.data(), sizeof(int32_t) * nE);
}

}  // extern "C"
