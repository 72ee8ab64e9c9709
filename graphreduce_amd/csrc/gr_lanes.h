// Lane-parallel kernels shared by the batched SSSP (gr_batched.hip) and the
// query server (gr_serve.hip): a wave64 handles one destination vertex and
// lane s works for query s, so a source's 64 values (256 B) are read with one
// coalesced load.
//
// Vertices with more than kLaneHeavy in-edges would serialise one wave over
// all of their edges; they are skipped by the per-vertex kernels and handled
// by edge chunks of kLaneChunk in-edges spread over many waves:
//   SSSP: each chunk's per-lane minimum is merged with atomicMin, and the lane
//         is marked changed when it lowered the stored distance;
//   PPR:  each chunk writes its per-lane partial sum, and a second kernel adds
//         the partials of a vertex in chunk order (deterministic).
#ifndef GRAPHREDUCE_AMD_GR_LANES_H
#define GRAPHREDUCE_AMD_GR_LANES_H

#include <algorithm>
#include <vector>

#include "gr_device.h"

namespace gr {

constexpr int kLaneNT = 256;
constexpr int kLaneHeavy = 4096;
constexpr int kLaneBatch = 16;  // rows read at once by a wave summing lanes
constexpr int kLaneChunk = 4096;

// (vertex, first in-edge, end in-edge, index of the vertex's first chunk)
struct LaneChunk {
  int32_t v, b, e, first;
};

// A heavy vertex and the range of its chunks.
struct LaneHeavy {
  int32_t v, first, count;
};

// Chunks of the heavy vertices of a CSC (host side, from the offsets).
inline void build_lane_chunks(const std::vector<int32_t>& off_in, std::vector<LaneChunk>& chunks,
                              std::vector<LaneHeavy>& heavy) {
  chunks.clear();
  heavy.clear();
  const int32_t nV = (int32_t)off_in.size() - 1;
  for (int32_t v = 0; v < nV; ++v) {
    const int32_t b = off_in[v], e = off_in[v + 1];
    if (e - b <= kLaneHeavy) continue;
    const int32_t first = (int32_t)chunks.size();
    for (int32_t c = b; c < e; c += kLaneChunk)
      chunks.push_back(LaneChunk{v, c, c + kLaneChunk < e ? c + kLaneChunk : e, first});
    heavy.push_back(LaneHeavy{v, first, (int32_t)chunks.size() - first});
  }
}

// Rows per segment of the segment walks, and the most in-edges a segment
// walked by one wave may hold.
constexpr int kLaneSegRows = 16;
constexpr int kLaneSegEdges = 1024;

// First vertex (a multiple of kLaneSegRows) from which on every segment of
// kLaneSegRows vertices holds at most kLaneSegEdges in-edges and no heavy
// row: the tail of the degree order, which the relax kernels walk a segment
// per wave; the vertices before it are taken one per wave.
inline int32_t lane_seg_split(const std::vector<int32_t>& off_in) {
  const int32_t nV = (int32_t)off_in.size() - 1;
  int32_t split = (nV + kLaneSegRows - 1) / kLaneSegRows * kLaneSegRows;
  while (split > 0) {
    const int32_t a = split - kLaneSegRows, b = std::min<int32_t>(split, nV);
    bool light = off_in[b] - off_in[a] <= kLaneSegEdges;
    for (int32_t v = a; light && v < b; ++v) light = off_in[v + 1] - off_in[v] <= kLaneHeavy;
    if (!light) break;
    split = a;
  }
  return std::min(split, nV);
}

// lane_seg_split of a graph's CSC, computed once per graph (a pass over the
// vertices of the tail).
inline int32_t graph_lane_split(gr_graph* g) {
  if (g->lane_split < 0) g->lane_split = lane_seg_split(g->h_off_in);
  return g->lane_split;
}

// Device copies of the chunk tables of a graph.
struct LaneTables {
  LaneChunk* chunks = nullptr;
  LaneHeavy* heavy = nullptr;
  int n_chunks = 0, n_heavy = 0;
  ~LaneTables() {
    if (chunks) (void)hipFree(chunks);
    if (heavy) (void)hipFree(heavy);
  }
  int build(const std::vector<int32_t>& off_in) {
    std::vector<LaneChunk> c;
    std::vector<LaneHeavy> h;
    build_lane_chunks(off_in, c, h);
    n_chunks = (int)c.size();
    n_heavy = (int)h.size();
    GR_HIP(hipMalloc((void**)&chunks, sizeof(LaneChunk) * (c.size() + 1)));
    GR_HIP(hipMalloc((void**)&heavy, sizeof(LaneHeavy) * (h.size() + 1)));
    if (n_chunks)
      GR_HIP(hipMemcpy(chunks, c.data(), sizeof(LaneChunk) * c.size(), hipMemcpyHostToDevice));
    if (n_heavy)
      GR_HIP(hipMemcpy(heavy, h.data(), sizeof(LaneHeavy) * h.size(), hipMemcpyHostToDevice));
    return GR_OK;
  }
};

__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src) {
  return ((uint64_t)__shfl((uint32_t)(x >> 32), src, kWave) << 32) |
         (uint64_t)__shfl((uint32_t)x, src, kWave);
}

// Per-lane minimum of dist[u][lane] + w over the in-edges [b, e) (b and e
// wave-uniform) whose source changed for this lane; up to kLaneBatch changed
// sources' rows are read at once.
__device__ __forceinline__ int32_t lane_min_in_edges(int b, int e, const int32_t* __restrict__ srcs,
                                                     const int32_t* __restrict__ w_csc,
                                                     const int32_t* __restrict__ dist,
                                                     const uint64_t* __restrict__ chg_in,
                                                     int32_t best) {
  const int lane = lane_id();
  const uint64_t my_bit = 1ull << lane;
  for (int pb = b; pb < e; pb += kWave) {
    const int p = pb + lane;
    int u = 0, wt = 0;
    uint64_t cu = 0;
    if (p < e) {
      u = srcs[p];
      cu = chg_in[u];
      if (cu) wt = w_csc[p];
    }
    uint64_t todo = __ballot(cu != 0);
    while (todo) {
      int uu[kLaneBatch], ww[kLaneBatch];
      uint64_t cm[kLaneBatch];
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j) {
        const int k = todo ? __builtin_ctzll(todo) : 0;
        cm[j] = todo ? lane_value64(cu, k) : 0ull;
        uu[j] = lane_value(u, k);
        ww[j] = lane_value(wt, k);
        todo &= todo - 1;
      }
      int32_t du[kLaneBatch];
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j)
        du[j] = (cm[j] & my_bit) ? dist[(size_t)uu[j] * kWave + lane] : 0;
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j)
        if (cm[j] & my_bit) {
          const int32_t cand = du[j] + ww[j];
          if (cand < best) best = cand;
        }
    }
  }
  return best;
}

// Round control of the lane SSSP.  A round only needs the vertices with an
// in-neighbour whose mask changed in the previous round; when those sources
// have few out-edges, k_lane_sssp_mark flags their out-neighbours in `dirty`,
// k_lane_sssp_filter turns the flagged light vertices into a list (zeroing
// chg_out of every other vertex and clearing the light flags), the relax
// kernels visit only the listed vertices and the flagged heavy chunks, and
// k_lane_sssp_clear_heavy clears the heavy flags, so `dirty` is all zero
// between rounds.  A vertex outside the list has no changed in-neighbour and
// would compute nothing, so the filter changes no value.
struct LaneRound {
  const unsigned long long* prev_count;  // previous round's changed vertices (nullptr: none)
  const unsigned long long* prev_edges;  // ... and their out-edges
  unsigned long long* count;             // this round's, counted when not nullptr
  unsigned long long* edges;
  uint8_t* dirty;                        // [V] flags; nullptr: never filter
  int32_t* list;                         // [V] light vertices to visit (filtered rounds)
  unsigned int* list_len;                // zeroed by the caller before the filter kernel
  long long filter_max;                  // filter when prev_edges < filter_max ...
  int host_filter;                       // ... (-1), or as the host says (0 / 1)
};

// A round after a round without change: nothing can change, every kernel of
// the round returns at once (chg_out is not rewritten: the caller stops).
__device__ __forceinline__ bool lane_round_idle(const LaneRound& r) {
  return r.prev_count && *r.prev_count == 0;
}
__device__ __forceinline__ bool lane_round_filtered(const LaneRound& r) {
  if (!r.dirty) return false;
  if (r.host_filter >= 0) return r.host_filter != 0;
  return r.prev_count && r.prev_edges && *r.prev_count != 0 &&
         (long long)*r.prev_edges < r.filter_max;
}

// Flags the out-neighbours of every vertex whose mask is set in chg_in.  Light
// out-degrees (<= kLaneHeavy) are walked by one wave per vertex; heavier
// sources by one wave per out-edge chunk (out_chunks: LaneChunk over CSR).
static __global__ __launch_bounds__(kLaneNT) void k_lane_sssp_mark(
    int nV, const int32_t* __restrict__ off_out, const int32_t* __restrict__ dsts,
    const LaneChunk* __restrict__ out_chunks, int n_out_chunks,
    const uint64_t* __restrict__ chg_in, LaneRound rr) {
  if (lane_round_idle(rr) || !lane_round_filtered(rr)) return;
  const int lane = lane_id();
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int n_waves = gridDim.x * blockDim.x / kWave;
  for (int base = wave * kWave; base < nV; base += n_waves * kWave) {
    const int v = base + lane;
    const uint64_t m = v < nV ? chg_in[v] : 0ull;
    int b = 0, e = 0;
    if (m) {
      b = off_out[v];
      e = off_out[v + 1];
    }
    uint64_t todo = __ballot(m != 0 && e - b <= kLaneHeavy);
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int bj = __shfl(b, j, kWave), ej = __shfl(e, j, kWave);
      for (int p = bj + lane; p < ej; p += kWave) rr.dirty[dsts[p]] = 1;
    }
  }
  for (int c = wave; c < n_out_chunks; c += n_waves) {
    const LaneChunk ch = out_chunks[c];
    if (!chg_in[ch.v]) continue;
    for (int p = ch.b + lane; p < ch.e; p += kWave) rr.dirty[dsts[p]] = 1;
  }
}

// Filtered rounds: flagged light vertices -> list; chg_out = 0 and flag
// cleared elsewhere (heavy flags stay for k_lane_sssp_heavy).  A workgroup
// owns a contiguous range of vertices: it counts its list entries, takes their
// positions with ONE atomic (same-address atomics of every wave would
// serialise), then writes them in ascending order.
static __global__ __launch_bounds__(kLaneNT) void k_lane_sssp_filter(
    int nV, const int32_t* __restrict__ off_in, uint64_t* __restrict__ chg_out, LaneRound rr) {
  if (lane_round_idle(rr) || !lane_round_filtered(rr)) return;
  __shared__ unsigned int s_w[kLaneNT / kWave];
  __shared__ unsigned int s_base;
  const int lane = lane_id(), w = threadIdx.x / kWave;
  const int64_t per = ((int64_t)nV + gridDim.x - 1) / gridDim.x;
  const int64_t span = (per + kLaneNT - 1) / kLaneNT * kLaneNT;
  const int64_t lo64 = (int64_t)blockIdx.x * span;
  const int lo = (int)(lo64 < nV ? lo64 : nV);
  const int hi = (int)(lo64 + span < nV ? lo64 + span : nV);
  auto wanted = [&](int v) {
    return v < hi && rr.dirty[v] != 0 && off_in[v + 1] - off_in[v] <= kLaneHeavy;
  };
  unsigned int mine = 0;
  for (int v = lo + (int)threadIdx.x; v < hi; v += kLaneNT) mine += wanted(v) ? 1u : 0u;
  // block total
  unsigned int tot = mine;
  for (int o = kWave / 2; o > 0; o >>= 1) tot += __shfl_xor(tot, o, kWave);
  if (lane == 0) s_w[w] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int t = 0;
    for (int i = 0; i < kLaneNT / kWave; ++i) t += s_w[i];
    s_base = t ? atomicAdd(rr.list_len, t) : 0u;
  }
  __syncthreads();
  unsigned int at = s_base;
  for (int v0 = lo; v0 < hi; v0 += kLaneNT) {
    const int v = v0 + (int)threadIdx.x;
    bool take = false;
    if (v < hi) {
      const bool flagged = rr.dirty[v] != 0;
      const bool heavy = off_in[v + 1] - off_in[v] > kLaneHeavy;
      take = flagged && !heavy;
      if (take) rr.dirty[v] = 0;
      if (!take) chg_out[v] = 0;  // heavy: k_lane_sssp_heavy ORs its bits in
    }
    const uint64_t tk = __ballot(take);
    __syncthreads();  // s_w of the previous step has been read
    if (lane == 0) s_w[w] = (unsigned int)__popcll(tk);
    __syncthreads();
    unsigned int before = 0, step = 0;
    for (int i = 0; i < kLaneNT / kWave; ++i) {
      if (i < w) before += s_w[i];
      step += s_w[i];
    }
    if (take) rr.list[at + before + __popcll(tk & ((1ull << lane) - 1))] = v;
    at += step;
  }
}

// One relaxation round over the light vertices (in-degree <= kLaneHeavy), all
// of them or the filtered list.  chg_out[v] = lanes whose distance at v
// dropped (heavy vertices get 0 here and their bits from k_lane_sssp_heavy);
// *any |= lanes changed anywhere; rr.count / rr.edges += changed vertices and
// their out-edges.  In unfiltered rounds the vertices from seg_split on (the
// tail of the degree order: few in-edges each) are taken kLaneSegRows at a
// time: one walk over their CSC segment, closing a vertex when the walk
// passes its last in-edge, instead of a chain of reads per vertex.  Every loop
// bound and cross-lane value is read out of its lane as a scalar
// (lane_value), so each wave runs every iteration with all of its lanes.
static __global__ __launch_bounds__(kLaneNT) void k_lane_sssp_light(
    int nV, const int32_t* __restrict__ off_in, const int32_t* __restrict__ off_out,
    const int32_t* __restrict__ srcs, const int32_t* __restrict__ w_csc,
    int32_t* __restrict__ dist, const uint64_t* __restrict__ chg_in,
    uint64_t* __restrict__ chg_out, unsigned long long* __restrict__ any, LaneRound rr,
    int seg_split) {
  if (lane_round_idle(rr)) return;
  const bool filtered = lane_round_filtered(rr);
  const int n_items = filtered ? (int)*rr.list_len : seg_split;
  __shared__ unsigned long long s_acc[kLaneNT / kWave], s_cnt[kLaneNT / kWave],
      s_edg[kLaneNT / kWave];
  const int lane = lane_id();
  const int wpb = kLaneNT / kWave;
  const int n_waves = gridDim.x * wpb;
  const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (int)threadIdx.x / kWave);
  const uint64_t my_bit = 1ull << lane;
  uint64_t acc = 0;
  unsigned long long cnt = 0, edg = 0;
  for (int i = w0; i < n_items; i += n_waves) {
    const int v = filtered ? __builtin_amdgcn_readfirstlane(rr.list[i]) : i;
    const int b = __builtin_amdgcn_readfirstlane(off_in[v]);
    const int e = __builtin_amdgcn_readfirstlane(off_in[v + 1]);
    if (e - b > kLaneHeavy) {
      if (lane == 0) chg_out[v] = 0;
      continue;
    }
    const int32_t before = dist[(size_t)v * kWave + lane];
    const int32_t best = lane_min_in_edges(b, e, srcs, w_csc, dist, chg_in, before);
    const bool improved = best < before;
    if (improved) dist[(size_t)v * kWave + lane] = best;
    const uint64_t m = __ballot(improved);
    if (lane == 0) chg_out[v] = m;
    acc |= m;
    if (m) {
      cnt += 1;
      if (lane == 0 && rr.edges) edg += (unsigned long long)(off_out[v + 1] - off_out[v]);
    }
  }
  const int n_seg = filtered ? 0 : (nV - seg_split + kLaneSegRows - 1) / kLaneSegRows;
  for (int sg = w0; sg < n_seg; sg += n_waves) {
    const int v0 = seg_split + sg * kLaneSegRows;
    const int rows = min(kLaneSegRows, nV - v0);
    // lanes 0 .. rows hold the rows' CSC and CSR offsets
    const int off = lane <= rows ? off_in[v0 + lane] : 0;
    const int oo = (lane <= rows && rr.edges) ? off_out[v0 + lane] : 0;
    int32_t bef[kLaneSegRows];
#pragma unroll
    for (int r = 0; r < kLaneSegRows; ++r)
      bef[r] = r < rows ? dist[(size_t)(v0 + r) * kWave + lane] : 0;
    const int seg_b = lane_value(off, 0), seg_e = lane_value(off, rows);
    int r = 0;                 // vertex (within the segment) whose in-edges the walk is in
    int32_t best = INT_MAX;    // its candidates' minimum, this lane
    // closes vertex r: a lane improves where a candidate beat the distance
    auto close_row = [&]() {
      int32_t b0 = bef[0];
#pragma unroll
      for (int k = 1; k < kLaneSegRows; ++k)
        if (k == r) b0 = bef[k];
      const int v = v0 + r;
      const bool improved = best < b0;
      if (improved) dist[(size_t)v * kWave + lane] = best;
      const uint64_t m = __ballot(improved);
      if (lane == 0) chg_out[v] = m;
      acc |= m;
      if (m) {
        cnt += 1;
        if (lane == 0 && rr.edges)
          edg += (unsigned long long)(lane_value(oo, r + 1) - lane_value(oo, r));
      }
      best = INT_MAX;
      ++r;
    };
    for (int pb = seg_b; pb < seg_e; pb += kWave) {
      const int p = pb + lane;
      int u = 0, wt = 0;
      uint64_t cu = 0;
      if (p < seg_e) {
        u = srcs[p];
        cu = chg_in[u];
        if (cu) wt = w_csc[p];
      }
      uint64_t todo = __ballot(cu != 0);
      while (todo) {
        int kk[kLaneBatch], uu[kLaneBatch], ww[kLaneBatch];
        uint64_t cm[kLaneBatch];
        int nb = 0;
#pragma unroll
        for (int j = 0; j < kLaneBatch; ++j) {
          kk[j] = todo ? __builtin_ctzll(todo) : 0;
          cm[j] = todo ? lane_value64(cu, kk[j]) : 0ull;
          uu[j] = lane_value(u, kk[j]);
          ww[j] = lane_value(wt, kk[j]);
          if (todo) nb = j + 1;
          todo &= todo - 1;
        }
        int32_t du[kLaneBatch];
#pragma unroll
        for (int j = 0; j < kLaneBatch; ++j)
          du[j] = (cm[j] & my_bit) ? dist[(size_t)uu[j] * kWave + lane] : 0;
#pragma unroll
        for (int j = 0; j < kLaneBatch; ++j) {
          if (j < nb) {
            // close the vertices whose in-edges all lie before this edge (the
            // segment's last vertex ends past every edge: r < rows)
            while (lane_value(off, r + 1) <= pb + kk[j]) close_row();
            if (cm[j] & my_bit) {
              const int32_t cand = du[j] + ww[j];
              if (cand < best) best = cand;
            }
          }
        }
      }
    }
    while (r < rows) close_row();
  }
  // one atomic per counter and workgroup: per-wave atomics on these words
  // serialise over hundreds of thousands of waves
  if (lane == 0) {
    s_acc[threadIdx.x / kWave] = acc;
    s_cnt[threadIdx.x / kWave] = cnt;
    s_edg[threadIdx.x / kWave] = edg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, c = 0, d = 0;
    for (int i = 0; i < wpb; ++i) {
      a |= s_acc[i];
      c += s_cnt[i];
      d += s_edg[i];
    }
    if (a) atomicOr(any, a);
    if (c && rr.count) atomicAdd(rr.count, c);
    if (d && rr.edges) atomicAdd(rr.edges, d);
  }
}

// Heavy vertices: one wave per chunk, merged with atomicMin; in filtered
// rounds only the chunks of flagged vertices.
static __global__ __launch_bounds__(kLaneNT) void k_lane_sssp_heavy(
    const LaneChunk* __restrict__ chunks, int n_chunks, const int32_t* __restrict__ off_out,
    const int32_t* __restrict__ srcs, const int32_t* __restrict__ w_csc,
    int32_t* __restrict__ dist, const uint64_t* __restrict__ chg_in,
    uint64_t* __restrict__ chg_out, unsigned long long* __restrict__ any, LaneRound rr) {
  if (lane_round_idle(rr)) return;
  const bool filtered = lane_round_filtered(rr);
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < n_chunks; c += n_waves) {
    const LaneChunk ch = chunks[c];
    if (filtered && !rr.dirty[ch.v]) continue;
    const int32_t inf = 2147483647;
    const int32_t best = lane_min_in_edges(ch.b, ch.e, srcs, w_csc, dist, chg_in, inf);
    bool improved = false;
    if (best < inf) {
      const int32_t old = atomicMin(&dist[(size_t)ch.v * kWave + lane], best);
      improved = best < old;
    }
    const uint64_t m = __ballot(improved);
    if (lane == 0 && m) {
      const unsigned long long prev =
          atomicOr((unsigned long long*)&chg_out[ch.v], (unsigned long long)m);
      atomicOr(any, (unsigned long long)m);
      if (prev == 0) {
        if (rr.count) atomicAdd(rr.count, 1ull);
        if (rr.edges) atomicAdd(rr.edges, (unsigned long long)(off_out[ch.v + 1] - off_out[ch.v]));
      }
    }
  }
}

// Filtered rounds: clears the flags of the heavy vertices after their chunks ran.
static __global__ void k_lane_sssp_clear_heavy(const LaneHeavy* __restrict__ heavy, int n_heavy,
                                               LaneRound rr) {
  if (lane_round_idle(rr) || !lane_round_filtered(rr)) return;
  for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < n_heavy; h += gridDim.x * blockDim.x)
    rr.dirty[heavy[h].v] = 0;
}

// PPR partial sum of contrib[u][lane] over in-edges [b, e) whose source has
// mass for an active lane, in ascending edge order.
__device__ __forceinline__ float lane_sum_in_edges(int b, int e, uint64_t active,
                                                   const int32_t* __restrict__ srcs,
                                                   const float* __restrict__ contrib,
                                                   const uint64_t* __restrict__ nz_in) {
  // b and e are wave-uniform; up to kLaneBatch mass rows are read at once
  // and added in edge order
  const int lane = lane_id();
  const bool mine = (active >> lane) & 1ull;
  float acc = 0.0f;
  for (int pb = b; pb < e; pb += kWave) {
    const int p = pb + lane;
    int u = 0;
    uint64_t m = 0;
    if (p < e) {
      u = srcs[p];
      m = nz_in[u] & active;
    }
    uint64_t todo = __ballot(m != 0);
    while (todo) {
      int uu[kLaneBatch];
      int nb = 0;
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j) {
        uu[j] = lane_value(u, todo ? __builtin_ctzll(todo) : 0);
        if (todo) nb = j + 1;
        todo &= todo - 1;
      }
      float x[kLaneBatch];
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j)
        x[j] = (mine && j < nb) ? contrib[(size_t)uu[j] * kWave + lane] : 0.0f;
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j)
        if (mine && j < nb) acc = __fadd_rn(acc, x[j]);
    }
  }
  return acc;
}

}  // namespace gr

#endif
