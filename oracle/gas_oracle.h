// CPU oracle for the Gather-Apply-Scatter hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a restatement, in our own code, of the
// reference's sequential CPU engine GASEngineRef (refgas.h:30-230), of its
// CSR/CSC builders (util.cuh:48-121) and of the four vertex programs and their
// driver loops (bfs.cu, sssp.cu, pagerank.cu, connected_component.cu).  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
// the checker of the HIP engine and as the timed CPU baseline; the product
// never links it.
//
// Differences from the reference, on purpose:
//  - indSort (util.cuh:48-72) uses an unstable std::sort, so the order of
//    in-edges inside one CSC row is unspecified there; here the sort is stable
//    (edge-list order within a row).  Only PageRank's FP32 sum order depends
//    on it.
//  - The BFS level is passed to apply() as an argument instead of the global
//    g_iterationCount (bfs.cu:27, 63-67); the value is the same.
//
// Pinning (what checks this oracle, see tests/test_oracle.py): BFS depths
// against scipy breadth-first levels, SSSP distances against scipy Dijkstra,
// CC labels against scipy connected components (min id per component),
// PageRank against closed forms (cycles, stars, sources) and an FP64 sparse
// mat-vec of the same recurrence, iteration counts against their closed form
// for BFS, and brute-force enumeration on tiny graphs.
#ifndef GRAPHREDUCE_ORACLE_GAS_ORACLE_H
#define GRAPHREDUCE_ORACLE_GAS_ORACLE_H

#include <stdint.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <vector>

namespace oracle {

// CSR of (rows -> cols) from an edge list; within a row the edges keep their
// edge-list order.  perm[p] = edge-list index of position p (util.cuh:79-110).
inline void edge_list_to_csr(int32_t n_vertices, int64_t n_edges, const int32_t* rows,
                             const int32_t* cols, std::vector<int32_t>& offsets,
                             std::vector<int32_t>& adj, std::vector<int32_t>& perm) {
  offsets.assign((size_t)n_vertices + 1, 0);
  for (int64_t i = 0; i < n_edges; ++i) offsets[(size_t)rows[i] + 1]++;
  for (int32_t v = 0; v < n_vertices; ++v) offsets[v + 1] += offsets[v];
  std::vector<int32_t> fill(offsets.begin(), offsets.end() - 1);
  adj.resize((size_t)n_edges);
  perm.resize((size_t)n_edges);
  for (int64_t i = 0; i < n_edges; ++i) {
    int32_t p = fill[rows[i]]++;
    adj[p] = cols[i];
    perm[p] = (int32_t)i;
  }
}

struct Stats {
  int64_t iterations = 0;     // loop trips with a non-empty active set
  int64_t active = 0;         // sum of #A over gathers
  int64_t in_edges = 0;       // sum of E_in(A) over gathers
  int64_t changed = 0;        // sum of #C (apply returned true)
  int64_t out_edges = 0;      // sum of E_out(C)
};

// The engine of refgas.h:30-230: bulk-synchronous gather over CSC in-edges,
// apply, activation of CSR out-neighbours, ascending next active list.
template <class Program>
class Engine {
 public:
  typedef typename Program::VertexData VertexData;
  typedef typename Program::EdgeData EdgeData;
  typedef typename Program::GatherResult GatherResult;

  // refgas.h:81-113
  void setGraph(int32_t nV, VertexData* vd, int64_t nE, EdgeData* ed, const int32_t* srcs,
                const int32_t* dsts) {
    nV_ = nV;
    vd_ = vd;
    ed_ = ed;
    edge_list_to_csr(nV, nE, srcs, dsts, dst_off_, dsts_, csr_perm_);
    edge_list_to_csr(nV, nE, dsts, srcs, src_off_, srcs_, csc_perm_);
    prepare();
  }

  // Same engine over prebuilt CSR (out) and CSC (in) arrays; edge data is then
  // indexed by CSC position and CSR position directly.
  void setGraphCSR(int32_t nV, VertexData* vd, const int32_t* off_out, const int32_t* dsts,
                   const int32_t* off_in, const int32_t* srcs, EdgeData* ed_csc) {
    nV_ = nV;
    vd_ = vd;
    ed_ = ed_csc;
    const int64_t nE = off_out[nV];
    dst_off_.assign(off_out, off_out + nV + 1);
    dsts_.assign(dsts, dsts + nE);
    src_off_.assign(off_in, off_in + nV + 1);
    srcs_.assign(srcs, srcs + nE);
    csc_perm_.resize((size_t)nE);
    std::iota(csc_perm_.begin(), csc_perm_.end(), 0);
    csr_perm_.clear();
    prepare();
  }

  // refgas.h:128-133
  void setActive(int32_t start, int32_t end) {
    active_.clear();
    for (int32_t i = start; i < end; ++i) active_.push_back(i);
  }

  int32_t countActive() const { return (int32_t)active_.size(); }

  // refgas.h:143-160 (the flag is ignored, as there)
  void gather(bool = true) {
    stats.active += (int64_t)active_.size();
    for (size_t i = 0; i < active_.size(); ++i) {
      const int32_t dv = active_[i];
      GatherResult sum = Program::gatherZero;
      stats.in_edges += src_off_[dv + 1] - src_off_[dv];
      for (int32_t ie = src_off_[dv]; ie < src_off_[dv + 1]; ++ie) {
        const int32_t src = srcs_[ie];
        GatherResult tmp = Program::gatherMap(vd_ + dv, vd_ + src, edge(csc_perm_[ie]));
        sum = Program::gatherReduce(sum, tmp);
      }
      gathered_[i] = sum;
    }
  }

  // refgas.h:163-171
  void apply(int iteration = 0) {
    for (size_t i = 0; i < active_.size(); ++i)
      apply_ret_[i] = Program::apply(vd_ + active_[i], gathered_[i], iteration);
  }

  // refgas.h:175-200
  void scatterActivate(bool haveScatter = true) {
    std::fill(flags_.begin(), flags_.end(), 0);
    for (size_t i = 0; i < active_.size(); ++i) {
      if (!apply_ret_[i]) continue;
      const int32_t sv = active_[i];
      stats.changed += 1;
      stats.out_edges += dst_off_[sv + 1] - dst_off_[sv];
      for (int32_t ie = dst_off_[sv]; ie < dst_off_[sv + 1]; ++ie) {
        const int32_t dv = dsts_[ie];
        flags_[dv] = 1;
        if (haveScatter)
          Program::scatter(vd_ + sv, vd_ + dv, edge(csr_perm_.empty() ? ie : csr_perm_[ie]));
      }
    }
  }

  // refgas.h:205-214
  int32_t nextIter() {
    if (!active_.empty()) stats.iterations += 1;
    active_.clear();
    for (int32_t i = 0; i < nV_; ++i)
      if (flags_[i]) active_.push_back(i);
    return countActive();
  }

  // refgas.h:220-229
  int32_t run() {
    int32_t it = 0;
    while (countActive()) {
      gather();
      apply(it);
      scatterActivate();
      nextIter();
      ++it;
    }
    return it;
  }

  const std::vector<int32_t>& outOffsets() const { return dst_off_; }

  Stats stats;

 private:
  void prepare() {
    active_.clear();
    active_.reserve((size_t)nV_);
    apply_ret_.assign((size_t)nV_, 0);
    flags_.assign((size_t)nV_, 0);
    gathered_.assign((size_t)nV_, GatherResult());
  }
  EdgeData* edge(int32_t i) { return ed_ ? ed_ + i : nullptr; }

  int32_t nV_ = 0;
  VertexData* vd_ = nullptr;
  EdgeData* ed_ = nullptr;
  std::vector<int32_t> src_off_, srcs_, csc_perm_;
  std::vector<int32_t> dst_off_, dsts_, csr_perm_;
  std::vector<int32_t> active_;
  std::vector<int32_t> apply_ret_;
  std::vector<char> flags_;
  std::vector<GatherResult> gathered_;
};

// ------------------------------------------------------- vertex programs

// bfs.cu:31-80
struct BFS {
  struct VertexData {
    int depth;
  };
  struct EdgeData {};
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX - 1;
  static int gatherReduce(const int&, const int&) { return 0; }
  static int gatherMap(const VertexData*, const VertexData*, const EdgeData*) { return 0; }
  static bool apply(VertexData* vert, int, int iteration) {
    if (vert->depth == -1) {
      vert->depth = iteration;
      return true;
    }
    return false;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// sssp.cu:25-66
struct SSSP {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static constexpr int maxLength = 100000;
  static constexpr int gatherZero = INT_MAX - maxLength;
  static int gatherReduce(const int& a, const int& b) { return std::min(a, b); }
  static int gatherMap(const VertexData*, const VertexData* srcDist, const EdgeData* len) {
    return *srcDist + *len;
  }
  static bool apply(VertexData* cur, GatherResult dist, int) {
    bool changed = dist < *cur;
    *cur = std::min(*cur, dist);
    return changed;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// A program outside the reference's four, for the open Program interface
// (gpugas.h:1025-1027, refgas.h:30-31): widest (maximum-bottleneck) paths.
// A vertex holds the largest capacity of a path from the source (INT_MAX at
// the source, 0 unreached); gatherMap takes the smaller of the source's width
// and the edge's capacity, gatherReduce the larger of two, apply keeps an
// increase.
struct WidestPath {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static constexpr int gatherZero = 0;
  static int gatherReduce(const int& a, const int& b) { return std::max(a, b); }
  static int gatherMap(const VertexData*, const VertexData* srcWidth, const EdgeData* cap) {
    return std::min(*srcWidth, *cap);
  }
  static bool apply(VertexData* cur, GatherResult width, int) {
    if (width > *cur) {
      *cur = width;
      return true;
    }
    return false;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// pagerank.cu:27-73.  Real = float is the reference program; Real = double
// runs the same activation-driven loop in FP64 (used to bound the FP32 sum
// error on vertices with very large in-degree).
template <class Real>
struct PageRankT {
  static constexpr Real pageConst = (Real)0.15f;
  static constexpr Real tol = (Real)0.01f;
  struct VertexData {
    Real rank;
    int numOutEdges;
  };
  struct EdgeData {};
  typedef Real GatherResult;
  static constexpr Real gatherZero = 0;
  static Real gatherMap(const VertexData*, const VertexData* src, const EdgeData*) {
    return src->rank / src->numOutEdges;
  }
  static Real gatherReduce(const Real& a, const Real& b) { return a + b; }
  static bool apply(VertexData* v, const Real& g, int) {
    Real newRank = pageConst + ((Real)1 - pageConst) * g;
    bool ret = std::fabs(newRank - v->rank) >= tol;
    v->rank = newRank;
    return ret;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};
typedef PageRankT<float> PageRank;
typedef PageRankT<double> PageRank64;

// connected_component.cu:25-63
struct CC {
  typedef int VertexData;
  struct EdgeData {};
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX;
  static int gatherReduce(const int& a, const int& b) { return std::min(a, b); }
  static int gatherMap(const VertexData*, const VertexData* src, const EdgeData*) { return *src; }
  static bool apply(VertexData* cur, GatherResult label, int) {
    bool changed = label < *cur;
    *cur = std::min(*cur, label);
    return changed;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// SSSP on congesting links (the library's GR_PROGRAM_SSSP_CONGESTION): the
// SSSP gather/apply, and a scatter that adds 1 to an out-edge of an improved
// vertex whose target is still farther away (called by scatterActivate as
// refgas.h:192-196 does).
struct SSSPCongestion {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX - 100000;
  static int gatherReduce(const int& a, const int& b) { return std::min(a, b); }
  static int gatherMap(const VertexData*, const VertexData* src, const EdgeData* w) {
    return *src + *w;
  }
  static bool apply(VertexData* cur, GatherResult d, int) {
    bool changed = d < *cur;
    *cur = std::min(*cur, d);
    return changed;
  }
  static void scatter(const VertexData* src, const VertexData* dst, EdgeData* w) {
    if (*dst > *src) *w += 1;
  }
};

}  // namespace oracle

#endif
