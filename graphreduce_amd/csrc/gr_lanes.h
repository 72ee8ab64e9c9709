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

#include <vector>

#include "gr_device.h"

namespace gr {

constexpr int kLaneNT = 256;
constexpr int kLaneHeavy = 4096;
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

// Per-lane minimum of dist[u][lane] + w over the in-edges [b, e) whose source
// changed for this lane; up to four changed edges are loaded at once.
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
      int k[4], uu[4], ww[4];
      uint64_t cm[4];
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        k[j] = todo ? __builtin_ctzll(todo) : 0;
        if (todo) {
          todo &= todo - 1;
          cnt = j + 1;
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        uu[j] = __shfl(u, k[j], kWave);
        ww[j] = __shfl(wt, k[j], kWave);
        cm[j] = shfl_u64(cu, k[j]);
      }
      int32_t du[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        du[j] = (j < cnt && (cm[j] & my_bit)) ? dist[(size_t)uu[j] * kWave + lane] : 0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j < cnt && (cm[j] & my_bit)) {
          const int32_t cand = du[j] + ww[j];
          if (cand < best) best = cand;
        }
    }
  }
  return best;
}

// One relaxation round over the light vertices (in-degree <= kLaneHeavy).
// chg_out[v] = lanes whose distance at v dropped; heavy vertices get 0 here
// and their bits from k_lane_sssp_heavy.  *any |= lanes changed anywhere.
// prev_count (may be nullptr): the previous round's changed count; when it is
// 0 nothing can change in this round and the kernel returns at once (chg_out
// is then not rewritten: the caller stops after such a round).
static __global__ __launch_bounds__(kLaneNT) void k_lane_sssp_light(
    int nV, const int32_t* __restrict__ off_in, const int32_t* __restrict__ srcs,
    const int32_t* __restrict__ w_csc, int32_t* __restrict__ dist,
    const uint64_t* __restrict__ chg_in, uint64_t* __restrict__ chg_out,
    unsigned long long* __restrict__ any, unsigned long long* __restrict__ count,
    const unsigned long long* __restrict__ prev_count) {
  if (prev_count && *prev_count == 0) return;
  __shared__ unsigned long long s_acc[kLaneNT / kWave], s_cnt[kLaneNT / kWave];
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  uint64_t acc = 0;
  unsigned long long cnt = 0;
  for (int v = blockIdx.x * wpb + threadIdx.x / kWave; v < nV; v += n_waves) {
    const int b = off_in[v], e = off_in[v + 1];
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
    cnt += m ? 1 : 0;
  }
  // one atomic pair per workgroup: per-wave atomics on these two words
  // serialise over hundreds of thousands of waves
  if (lane == 0) {
    s_acc[threadIdx.x / kWave] = acc;
    s_cnt[threadIdx.x / kWave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, c = 0;
    for (int i = 0; i < wpb; ++i) {
      a |= s_acc[i];
      c += s_cnt[i];
    }
    if (a) atomicOr(any, a);
    if (c && count) atomicAdd(count, c);
  }
}

// Heavy vertices: one wave per chunk, merged with atomicMin.
static __global__ __launch_bounds__(kLaneNT) void k_lane_sssp_heavy(
    const LaneChunk* __restrict__ chunks, int n_chunks, const int32_t* __restrict__ srcs,
    const int32_t* __restrict__ w_csc, int32_t* __restrict__ dist,
    const uint64_t* __restrict__ chg_in, uint64_t* __restrict__ chg_out,
    unsigned long long* __restrict__ any, unsigned long long* __restrict__ count,
    const unsigned long long* __restrict__ prev_count) {
  if (prev_count && *prev_count == 0) return;
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < n_chunks; c += n_waves) {
    const LaneChunk ch = chunks[c];
    const int32_t inf = 2147483647;
    const int32_t best = lane_min_in_edges(ch.b, ch.e, srcs, w_csc, dist, chg_in, inf);
    bool improved = false;
    if (best < inf) {
      const int32_t old = atomicMin(&dist[(size_t)ch.v * kWave + lane], best);
      improved = best < old;
    }
    const uint64_t m = __ballot(improved);
    if (lane == 0 && m) {
      const unsigned long long prev = atomicOr((unsigned long long*)&chg_out[ch.v], (unsigned long long)m);
      atomicOr(any, (unsigned long long)m);
      if (count && prev == 0) atomicAdd(count, 1ull);
    }
  }
}

// PPR partial sum of contrib[u][lane] over in-edges [b, e) whose source has
// mass for an active lane, in ascending edge order.
__device__ __forceinline__ float lane_sum_in_edges(int b, int e, uint64_t active,
                                                   const int32_t* __restrict__ srcs,
                                                   const float* __restrict__ contrib,
                                                   const uint64_t* __restrict__ nz_in) {
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
      int k[4];
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        k[j] = todo ? __builtin_ctzll(todo) : 0;
        if (todo) {
          todo &= todo - 1;
          cnt = j + 1;
        }
      }
      int uu[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) uu[j] = __shfl(u, k[j], kWave);
      float x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[j] = (mine && j < cnt) ? contrib[(size_t)uu[j] * kWave + lane] : 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (mine && j < cnt) acc = __fadd_rn(acc, x[j]);
    }
  }
  return acc;
}

}  // namespace gr

#endif
