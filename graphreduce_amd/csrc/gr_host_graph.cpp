// Host-side synthetic graph generators (RMAT, 2-D grid) and hash weights.
//
// These produce the inputs that the benchmark configurations name: RMAT with
// quadrant probabilities (a, b, c, d), self-loops and duplicate arcs removed,
// optional symmetrisation; a 4-neighbour grid with a seeded fraction of its
// edges deleted; uniform integer edge weights.  Output edge lists are sorted
// by (src, dst).  The device generator in gr_graph.hip draws the same numbers
// (gr_hash.h), so both sides agree bit for bit.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "gr_hash.h"
#include "gr_internal.h"

namespace gr {

namespace {

unsigned host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  if (n > 16) n = 16;
  return n;
}

template <class F>
void parallel_for(int64_t n, F&& f) {
  unsigned nt = host_threads();
  if (n < (int64_t)1 << 16 || nt == 1) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> ts;
  int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    int64_t b = (int64_t)t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& t : ts) t.join();
}

// LSD radix sort of keys that use the low `bits` bits, 16 bits per pass.
void radix_sort_u64(std::vector<uint64_t>& keys, int bits) {
  std::vector<uint64_t> tmp(keys.size());
  for (int shift = 0; shift < bits; shift += 16) {
    std::vector<int64_t> count(65537, 0);
    for (uint64_t k : keys) count[((k >> shift) & 0xFFFF) + 1]++;
    for (int i = 0; i < 65536; ++i) count[i + 1] += count[i];
    for (uint64_t k : keys) tmp[count[(k >> shift) & 0xFFFF]++] = k;
    keys.swap(tmp);
  }
}

// Sort, drop self-loops (src == dst) and duplicates; keys are src<<sh | dst.
void sort_unique(std::vector<uint64_t>& keys, int sh) {
  radix_sort_u64(keys, 2 * sh);
  const uint64_t mask = (sh == 64) ? ~0ULL : ((1ULL << sh) - 1);
  size_t w = 0;
  for (size_t i = 0; i < keys.size(); ++i) {
    uint64_t k = keys[i];
    if ((k >> sh) == (k & mask)) continue;
    if (w > 0 && keys[w - 1] == k) continue;
    keys[w++] = k;
  }
  keys.resize(w);
}

}  // namespace

int64_t host_rmat_keys(int scale, int edge_factor, double a, double b, double c,
                       uint64_t seed, int symmetrize, std::vector<uint64_t>& keys) {
  const int64_t m = (int64_t)edge_factor << scale;
  const RmatThresholds t = rmat_thresholds(a, b, c);
  const uint64_t base = stream_base(seed, kStreamRmat);
  keys.resize(symmetrize ? 2 * m : m);
  uint64_t* kp = keys.data();
  parallel_for(m, [&](int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) {
      uint32_t s, d;
      rmat_edge(base, (uint64_t)k, scale, t, &s, &d);
      kp[k] = ((uint64_t)s << scale) | d;
      if (symmetrize) kp[m + k] = ((uint64_t)d << scale) | s;
    }
  });
  sort_unique(keys, scale);
  return (int64_t)keys.size();
}

}  // namespace gr

using namespace gr;

extern "C" int gr_host_rmat_edges(int scale, int edge_factor, double a, double b, double c,
                                  uint64_t seed, int symmetrize, int32_t* srcs,
                                  int32_t* dsts, int64_t capacity, int64_t* n_edges) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || !n_edges || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return set_error(GR_ERR_INVALID, "gr_host_rmat_edges: bad arguments");
  if ((int64_t)edge_factor << scale > ((int64_t)1 << 31) / (symmetrize ? 2 : 1))
    return set_error(GR_ERR_INVALID, "gr_host_rmat_edges: edge count exceeds int32");
  std::vector<uint64_t> keys;
  int64_t n = host_rmat_keys(scale, edge_factor, a, b, c, seed, symmetrize, keys);
  *n_edges = n;
  if (!srcs && !dsts) return GR_OK;
  if (!srcs || !dsts || capacity < n)
    return set_error(GR_ERR_INVALID, "gr_host_rmat_edges: output capacity too small");
  const uint64_t mask = (1ULL << scale) - 1;
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      srcs[i] = (int32_t)(keys[i] >> scale);
      dsts[i] = (int32_t)(keys[i] & mask);
    }
  });
  return GR_OK;
}

namespace gr {

// Undirected grid edge ids: horizontal (r,c)-(r,c+1) -> r*(side-1)+c;
// vertical (r,c)-(r+1,c) -> side*(side-1) + r*side + c.
int64_t host_grid_arcs(int side, double delete_frac, uint64_t seed, int32_t* srcs,
                       int32_t* dsts) {
  const uint64_t base = stream_base(seed, kStreamGrid);
  const double thr_d = delete_frac * 18446744073709551616.0;
  const uint64_t thr = thr_d >= 18446744073709551615.0 ? ~0ULL : (uint64_t)thr_d;
  const int64_t nh = (int64_t)side * (side - 1);
  auto h_id = [&](int r, int c) { return (uint64_t)((int64_t)r * (side - 1) + c); };
  auto v_id = [&](int r, int c) { return (uint64_t)(nh + (int64_t)r * side + c); };
  int64_t w = 0;
  for (int r = 0; r < side; ++r) {
    for (int c = 0; c < side; ++c) {
      const int32_t v = r * side + c;
      // neighbours in ascending id order: up, left, right, down
      if (r > 0 && grid_keep(base, v_id(r - 1, c), thr)) {
        if (srcs) { srcs[w] = v; dsts[w] = v - side; }
        ++w;
      }
      if (c > 0 && grid_keep(base, h_id(r, c - 1), thr)) {
        if (srcs) { srcs[w] = v; dsts[w] = v - 1; }
        ++w;
      }
      if (c + 1 < side && grid_keep(base, h_id(r, c), thr)) {
        if (srcs) { srcs[w] = v; dsts[w] = v + 1; }
        ++w;
      }
      if (r + 1 < side && grid_keep(base, v_id(r, c), thr)) {
        if (srcs) { srcs[w] = v; dsts[w] = v + side; }
        ++w;
      }
    }
  }
  return w;
}

}  // namespace gr

extern "C" int gr_host_grid_edges(int side, double delete_frac, uint64_t seed, int32_t* srcs,
                                  int32_t* dsts, int64_t capacity, int64_t* n_edges) {
  if (side < 1 || side > 32768 || !n_edges || delete_frac < 0 || delete_frac > 1)
    return set_error(GR_ERR_INVALID, "gr_host_grid_edges: bad arguments");
  int64_t n = host_grid_arcs(side, delete_frac, seed, nullptr, nullptr);
  *n_edges = n;
  if (!srcs && !dsts) return GR_OK;
  if (!srcs || !dsts || capacity < n)
    return set_error(GR_ERR_INVALID, "gr_host_grid_edges: output capacity too small");
  host_grid_arcs(side, delete_frac, seed, srcs, dsts);
  return GR_OK;
}

extern "C" int gr_host_hash_weights(uint64_t seed, int64_t n, int32_t lo, int span_log2,
                                    int32_t* weights) {
  if (n < 0 || !weights || span_log2 < 1 || span_log2 > 31)
    return set_error(GR_ERR_INVALID, "gr_host_hash_weights: bad arguments");
  const uint64_t base = stream_base(seed, kStreamWeight);
  parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      weights[i] = (int32_t)hash_weight(base, (uint64_t)i, (uint32_t)lo, span_log2);
  });
  return GR_OK;
}
