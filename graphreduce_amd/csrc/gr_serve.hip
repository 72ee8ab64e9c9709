// Continuous-batching query server: BFS, SSSP and personalized-PageRank
// queries on one device-resident graph.
//
// Each query kind has a pool of up to 64 lanes whose state stays resident in
// HBM between steps (one bit per lane for BFS masks, one lane of a wave64 for
// SSSP distances and PPR ranks).  Per pool, a host thread runs the loop:
//   1. admit queued requests into free lanes, as long as the pool has a free
//      lane and the device-memory budget covers the lane's bytes;
//   2. run ONE step (a BFS level, an SSSP relaxation round, a PPR power
//      iteration) of the pool's active lanes; every pool has its own host
//      thread and HIP stream, so the three kinds overlap on the device and
//      each advances at its own rate;
//   3. retire the lanes that finished: the lane's result row is reduced to a
//      64-bit checksum on the device and, when the request asked for it,
//      copied to the host.
// New requests therefore join while others are mid-flight ("continuous
// batching"), and every lane computes exactly what a solo run computes:
// BFS levels and SSSP distances are schedule-independent, and a PPR lane's
// arithmetic never reads another lane, in a fixed order.
//
// Latency = completion time - submit time (host steady clock); throughput and
// p50/p99 are reported by gr_server_get_stats.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gr_device.h"
#include "gr_gather.h"
#include "gr_hash.h"
#include "gr_lanes.h"

namespace gr {

constexpr int kSrvNT = 256;
constexpr int kLanes = 64;
constexpr int kSrvHeavy = 64;
constexpr int kPprSegEdges = 4096;  // in-edges of a 64-vertex segment the PPR walk takes

// ------------------------------------------------------------------ kernels

// Clears the lanes in `mask` from the BFS masks.
__global__ void k_srv_bfs_clear(int nV, uint64_t keep, uint64_t* __restrict__ visit,
                                uint64_t* __restrict__ frontier, uint64_t* __restrict__ next) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
    visit[v] &= keep;
    frontier[v] &= keep;
    next[v] &= keep;
  }
}

struct LaneSeeds {
  int32_t vertex[kLanes];
  int32_t start[kLanes];  // BFS: pool step at which the lane was admitted
};

__global__ void k_srv_bfs_seed(uint64_t mask, LaneSeeds seeds, int nV, uint64_t* __restrict__ visit,
                               uint64_t* __restrict__ frontier, int32_t* __restrict__ depth) {
  const int s = threadIdx.x;
  if (s >= kLanes || !((mask >> s) & 1ull)) return;
  const int v = seeds.vertex[s];
  atomicOr((unsigned long long*)&visit[v], 1ull << s);
  atomicOr((unsigned long long*)&frontier[v], 1ull << s);
  depth[(size_t)v * kLanes + s] = 0;
}

// BFS: depth column of the admitted lanes = -1 (vertex-major [V][64]).
__global__ void k_srv_bfs_reset_depth(int nV, uint64_t mask, int32_t* __restrict__ depth) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const bool mine = (mask >> lane) & 1ull;
  for (int v = blockIdx.x * wpb + threadIdx.x / kWave; v < nV; v += gridDim.x * wpb)
    if (mine) depth[(size_t)v * kLanes + lane] = -1;
}

// Expand one level for every active lane (push when the frontier's out-edges
// are few, pull otherwise), as in k_msbfs.
__global__ __launch_bounds__(kSrvNT) void k_srv_bfs_expand(
    int nV, int pull, uint64_t active, const int32_t* __restrict__ off_out,
    const int32_t* __restrict__ dsts, const int32_t* __restrict__ off_in,
    const int32_t* __restrict__ srcs, const uint64_t* __restrict__ visit,
    const uint64_t* __restrict__ frontier, uint64_t* __restrict__ next) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  for (int base = (blockIdx.x * wpb + threadIdx.x / kWave) * kWave; base < nV;
       base += n_waves * kWave) {
    const int v = base + lane;
    const bool valid = v < nV;
    if (!pull) {
      uint64_t f = valid ? frontier[v] : 0ull;
      int b = 0, e = 0;
      if (f) {
        b = off_out[v];
        e = off_out[v + 1];
        if (e - b > kLaneHeavy) f = 0;  // chunked: k_srv_bfs_push_chunks
      }
      const bool light = (e - b) <= kSrvHeavy;
      if (f && light)
        for (int p = b; p < e; ++p) {
          const int u = dsts[p];
          const uint64_t want = f & ~visit[u] & ~next[u];  // an atomic only for new bits
          if (want) atomicOr((unsigned long long*)&next[u], (unsigned long long)want);
        }
      uint64_t heavy = __ballot(f && !light);
      while (heavy) {
        const int hl = __builtin_ctzll(heavy);
        heavy &= heavy - 1;
        const int hb = __shfl(b, hl, kWave), he = __shfl(e, hl, kWave);
        const uint64_t hf = ((uint64_t)__shfl((uint32_t)(f >> 32), hl, kWave) << 32) |
                            __shfl((uint32_t)f, hl, kWave);
        for (int p = hb + lane; p < he; p += kWave) {
          const int u = dsts[p];
          const uint64_t want = hf & ~visit[u] & ~next[u];
          if (want) atomicOr((unsigned long long*)&next[u], (unsigned long long)want);
        }
      }
    } else {
      uint64_t need = valid ? (active & ~visit[v]) : 0ull;
      int b = 0, e = 0;
      if (need) {
        b = off_in[v];
        e = off_in[v + 1];
        if (e - b > kLaneHeavy) need = 0;  // chunked: k_srv_bfs_pull_chunks (atomicOr)
      }
      const bool light = (e - b) <= kSrvHeavy;
      if (need && light) {
        uint64_t acc = 0;
        for (int p = b; p < e; ++p) {
          acc |= frontier[srcs[p]];
          if ((acc & need) == need) break;
        }
        next[v] = acc & need;
      }
      uint64_t heavy = __ballot(need && !light);
      while (heavy) {
        const int hl = __builtin_ctzll(heavy);
        heavy &= heavy - 1;
        const int hb = __shfl(b, hl, kWave), he = __shfl(e, hl, kWave);
        const uint64_t hn = ((uint64_t)__shfl((uint32_t)(need >> 32), hl, kWave) << 32) |
                            __shfl((uint32_t)need, hl, kWave);
        uint64_t hacc = 0;
        for (int pb = hb; pb < he; pb += kWave) {
          const int p = pb + lane;
          uint64_t m = p < he ? frontier[srcs[p]] : 0ull;
          uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
#pragma unroll
          for (int off = kWave / 2; off > 0; off >>= 1) {
            lo |= __shfl_xor(lo, off, kWave);
            hi |= __shfl_xor(hi, off, kWave);
          }
          hacc |= ((uint64_t)hi << 32) | lo;
          if ((hacc & hn) == hn) break;
        }
        if (lane == hl) next[v] = hacc & hn;
      }
    }
  }
}

struct SrvCounters {
  unsigned long long frontier_or;     // lanes with a non-empty frontier after the step
  unsigned long long frontier_edges;  // out-edges of the new frontier
  unsigned long long changed;         // SSSP: vertices changed by the step
  unsigned int list_len;              // SSSP: filtered-list length of the step
  unsigned int pad;
  unsigned long long checksum[kLanes];
};

// Vertices with more than kLaneHeavy edges: chunks of kLaneChunk edges spread
// over all waves (one wave per hub would be the critical path of a step).
__global__ __launch_bounds__(kSrvNT) void k_srv_bfs_push_chunks(
    const LaneChunk* __restrict__ chunks, int n_chunks, const int32_t* __restrict__ dsts,
    const uint64_t* __restrict__ visit, const uint64_t* __restrict__ frontier,
    uint64_t* __restrict__ next) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < n_chunks; c += gridDim.x * wpb) {
    const LaneChunk ch = chunks[c];
    const uint64_t f = frontier[ch.v];
    if (!f) continue;
    for (int p = ch.b + lane; p < ch.e; p += kWave) {
      const int u = dsts[p];
      const uint64_t want = f & ~visit[u] & ~next[u];
      if (want) atomicOr((unsigned long long*)&next[u], (unsigned long long)want);
    }
  }
}

__global__ __launch_bounds__(kSrvNT) void k_srv_bfs_pull_chunks(
    const LaneChunk* __restrict__ chunks, int n_chunks, uint64_t active,
    const int32_t* __restrict__ srcs, const uint64_t* __restrict__ visit,
    const uint64_t* __restrict__ frontier, uint64_t* __restrict__ next) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < n_chunks; c += gridDim.x * wpb) {
    const LaneChunk ch = chunks[c];
    const uint64_t need = active & ~visit[ch.v];
    if (!need) continue;
    uint64_t acc = 0;
    for (int pb = ch.b; pb < ch.e; pb += kWave) {  // wave-uniform trip count
      const int p = pb + lane;
      uint64_t m = p < ch.e ? frontier[srcs[p]] : 0ull;
      uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        lo |= __shfl_xor(lo, off, kWave);
        hi |= __shfl_xor(hi, off, kWave);
      }
      acc |= ((uint64_t)hi << 32) | lo;
      if ((acc & need) == need) break;
    }
    if (lane == 0 && (acc & need))
      atomicOr((unsigned long long*)&next[ch.v], (unsigned long long)(acc & need));
  }
}

// New (lane, vertex) pairs of the level: visit/frontier, depth (vertex-major
// [V][64], one masked row store per vertex with new lanes), and the step's
// counters summed per workgroup (one atomic pair per workgroup).
__global__ __launch_bounds__(kSrvNT) void k_srv_bfs_update(
    int nV, int step, uint64_t active, LaneSeeds seeds, const int32_t* __restrict__ off_out,
    uint64_t* __restrict__ visit, uint64_t* __restrict__ frontier, uint64_t* __restrict__ next,
    int32_t* __restrict__ depth, SrvCounters* __restrict__ ctr) {
  __shared__ unsigned long long s_or[kSrvNT / kWave], s_fe[kSrvNT / kWave];
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  const int my_depth_base = step + 1 - seeds.start[lane];  // depth of this lane's new vertices
  uint64_t f_or = 0;
  unsigned long long fe = 0;
  for (int base = (blockIdx.x * wpb + threadIdx.x / kWave) * kWave; base < nV;
       base += n_waves * kWave) {
    const int v = base + lane;
    uint64_t nw = 0;
    if (v < nV) {
      const uint64_t vis = visit[v];
      const uint64_t nxt = next[v];
      nw = nxt & ~vis & active;
      if (nxt) next[v] = 0;
      // lanes that are not active keep no frontier
      if (nw || frontier[v]) frontier[v] = nw;
      if (nw) {
        visit[v] = vis | nw;
        f_or |= nw;
        fe += (unsigned long long)(off_out[v + 1] - off_out[v]);
      }
    }
    uint64_t rows = __ballot(nw != 0);
    while (rows) {
      const int r = __builtin_ctzll(rows);
      rows &= rows - 1;
      const uint32_t lo = __shfl((uint32_t)nw, r, kWave);
      const uint32_t hi = __shfl((uint32_t)(nw >> 32), r, kWave);
      const uint64_t m = ((uint64_t)hi << 32) | lo;
      if ((m >> lane) & 1ull) depth[(size_t)(base + r) * kLanes + lane] = my_depth_base;
    }
  }
  uint32_t lo = (uint32_t)f_or, hi = (uint32_t)(f_or >> 32);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    lo |= __shfl_xor(lo, off, kWave);
    hi |= __shfl_xor(hi, off, kWave);
    fe += __shfl_xor(fe, off, kWave);
  }
  if (lane == 0) {
    s_or[threadIdx.x / kWave] = ((uint64_t)hi << 32) | lo;
    s_fe[threadIdx.x / kWave] = fe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long o = 0, e = 0;
    for (int i = 0; i < wpb; ++i) {
      o |= s_or[i];
      e += s_fe[i];
    }
    if (o) atomicOr(&ctr->frontier_or, o);
    if (e) atomicAdd(&ctr->frontier_edges, e);
  }
}

// SSSP: reset the admitted lanes' distance column and changed bits.
__global__ void k_srv_sssp_reset(int nV, uint64_t mask, int32_t inf, int32_t* __restrict__ dist,
                                 uint64_t* __restrict__ chg_a, uint64_t* __restrict__ chg_b) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const bool mine = (mask >> lane) & 1ull;
  for (int v = blockIdx.x * wpb + threadIdx.x / kWave; v < nV; v += gridDim.x * wpb) {
    if (mine) dist[(size_t)v * kLanes + lane] = inf;
    if (lane == 0) {
      chg_a[v] &= ~mask;
      chg_b[v] &= ~mask;
    }
  }
}

__global__ void k_srv_sssp_seed(uint64_t mask, LaneSeeds seeds, int32_t* __restrict__ dist,
                                uint64_t* __restrict__ chg) {
  const int s = threadIdx.x;
  if (s >= kLanes || !((mask >> s) & 1ull)) return;
  const int v = seeds.vertex[s];
  dist[(size_t)v * kLanes + s] = 0;
  atomicOr((unsigned long long*)&chg[v], 1ull << s);
}

// PPR: reset admitted lanes (rank = e_seed).
__global__ void k_srv_ppr_reset(int nV, uint64_t mask, LaneSeeds seeds, float* __restrict__ rank,
                                uint64_t* __restrict__ nz) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const bool mine = (mask >> lane) & 1ull;
  for (int v = blockIdx.x * wpb + threadIdx.x / kWave; v < nV; v += gridDim.x * wpb) {
    const bool seed = mine && seeds.vertex[lane] == v;
    if (mine) rank[(size_t)v * kLanes + lane] = seed ? 1.0f : 0.0f;
    const uint64_t sm = __ballot(seed);
    if (lane == 0) nz[v] = (nz[v] & ~mask) | sm;
  }
}

// contrib[u][s] = rank[u][s] / outdeg[u] (0 without out-edges).  A wave takes
// 64 consecutive vertices: their out-degrees in one read (lane per vertex),
// then their rows kLaneBatch at a time with all reads in flight.  Every loop
// bound and cross-lane value is a scalar, so all lanes run every iteration.
__global__ __launch_bounds__(kSrvNT) void k_srv_ppr_contrib(int nV, uint64_t active,
                                                          const int32_t* __restrict__ off_out,
                                                          const float* __restrict__ rank,
                                                          float* __restrict__ contrib) {
  const int lane = lane_id();
  const int wpb = kSrvNT / kWave;
  const int n_waves = gridDim.x * wpb;
  const bool mine = (active >> lane) & 1ull;
  const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (int)threadIdx.x / kWave);
  for (int v0 = w0 * kWave; v0 < nV; v0 += n_waves * kWave) {
    const int cnt = min(kWave, nV - v0);
    const int deg = lane < cnt ? off_out[v0 + lane + 1] - off_out[v0 + lane] : 0;
    for (int r0 = 0; r0 < cnt; r0 += kLaneBatch) {
      float x[kLaneBatch];
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j)
        x[j] = (mine && r0 + j < cnt) ? rank[(size_t)(v0 + r0 + j) * kLanes + lane] : 0.0f;
#pragma unroll
      for (int j = 0; j < kLaneBatch; ++j) {
        if (r0 + j < cnt) {
          const int d = lane_value(deg, r0 + j);
          if (mine) contrib[(size_t)(v0 + r0 + j) * kLanes + lane] = d ? __fdiv_rn(x[j], (float)d) : 0.0f;
        }
      }
    }
  }
}

// r'[v][s] = 0.15 * [v == seed_s] + 0.85 * sum_{u->v} contrib[u][s], in-edge
// order fixed; sources without mass for any active lane are skipped (adding
// +0.0 changes no sum that starts at +0.0).
//
// k_srv_ppr_step_rows: one wave per vertex of [v_begin, v_end), the vertices
// of the degree order's head, whose 64-vertex segments hold many in-edges.
__global__ __launch_bounds__(kSrvNT) void k_srv_ppr_step_rows(
    int v_begin, int v_end, uint64_t active, LaneSeeds seeds, const int32_t* __restrict__ off_in,
    const int32_t* __restrict__ srcs, const float* __restrict__ contrib,
    const uint64_t* __restrict__ nz_in, uint64_t* __restrict__ nz_out, float* __restrict__ rank) {
  const int lane = lane_id();
  const int wpb = kSrvNT / kWave;
  const int n_waves = gridDim.x * wpb;
  const bool mine = (active >> lane) & 1ull;
  const int my_seed = seeds.vertex[lane];
  const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (int)threadIdx.x / kWave);
  for (int v = v_begin + w0; v < v_end; v += n_waves) {
    const int b = __builtin_amdgcn_readfirstlane(off_in[v]);
    const int e = __builtin_amdgcn_readfirstlane(off_in[v + 1]);
    if (e - b > kLaneHeavy) continue;  // k_srv_ppr_heavy_*
    const float acc = lane_sum_in_edges(b, e, active, srcs, contrib, nz_in);
    float nr = __fmul_rn(0.85f, acc);
    if (mine && my_seed == v) nr = __fadd_rn(nr, 0.15f);
    if (mine) rank[(size_t)v * kLanes + lane] = nr;
    const uint64_t nzm = __ballot(mine && nr != 0.0f);
    if (lane == 0) nz_out[v] = (nz_in[v] & ~active) | nzm;
  }
}

// k_srv_ppr_step: a wave takes 64 consecutive vertices from v_begin (a
// multiple of 64) on.  Without a heavy row among them it walks their CSC segment once,
// 64 in-edges per round, and adds the mass rows kLaneBatch at a time in edge
// order, closing a vertex's sum when the walk passes its last in-edge: the
// same additions in the same order as a walk per vertex, without a chain of
// reads per vertex (most vertices have a few in-edges or none).  A segment
// with a heavy row (summed by k_srv_ppr_heavy_*) takes its vertices one by
// one.  The segment's offsets live one per lane; every loop bound, row index
// and source id the loops use is read out of its lane as a scalar
// (lane_value), so the wave runs each iteration with all lanes: a lane-wise
// loop test would let the compiler run lanes through different iterations,
// in which a shuffle from a lane that is not running reads 0.
__global__ __launch_bounds__(kSrvNT) void k_srv_ppr_step(
    int v_begin, int nV, uint64_t active, LaneSeeds seeds, const int32_t* __restrict__ off_in,
    const int32_t* __restrict__ srcs, const float* __restrict__ contrib,
    const uint64_t* __restrict__ nz_in, uint64_t* __restrict__ nz_out, float* __restrict__ rank) {
  const int lane = lane_id();
  const int wpb = kSrvNT / kWave;
  const int n_waves = gridDim.x * wpb;
  const bool mine = (active >> lane) & 1ull;
  const int my_seed = seeds.vertex[lane];
  const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (int)threadIdx.x / kWave);
  for (int v0 = v_begin + w0 * kWave; v0 < nV; v0 += n_waves * kWave) {
    const int cnt = min(kWave, nV - v0);
    const int beg = lane < cnt ? off_in[v0 + lane] : 0;
    const int end = lane < cnt ? off_in[v0 + lane + 1] : 0;
    const uint64_t nzr = lane < cnt ? nz_in[v0 + lane] : 0ull;
    if (__ballot(lane < cnt && end - beg > kLaneHeavy)) {
      for (int r = 0; r < cnt; ++r) {
        const int v = v0 + r;
        const int b = lane_value(beg, r), e = lane_value(end, r);
        if (e - b > kLaneHeavy) continue;  // k_srv_ppr_heavy_*
        const float acc = lane_sum_in_edges(b, e, active, srcs, contrib, nz_in);
        float nr = __fmul_rn(0.85f, acc);
        if (mine && my_seed == v) nr = __fadd_rn(nr, 0.15f);
        if (mine) rank[(size_t)v * kLanes + lane] = nr;
        const uint64_t nzm = __ballot(mine && nr != 0.0f);
        if (lane == 0) nz_out[v] = (lane_value64(nzr, r) & ~active) | nzm;
      }
      continue;
    }
    const int seg_b = lane_value(beg, 0), seg_e = lane_value(end, cnt - 1);
    int r = 0;  // vertex (within the segment) whose in-edges the walk is in
    float acc = 0.0f;
    for (int pb = seg_b; pb < seg_e; pb += kWave) {
      const int p = pb + lane;
      int u = 0;
      uint64_t m = 0;
      if (p < seg_e) {
        u = srcs[p];
        m = nz_in[u] & active;
      }
      uint64_t todo = __ballot(m != 0);
      while (todo) {
        int k[kLaneBatch];
        int nb = 0;
#pragma unroll
        for (int j = 0; j < kLaneBatch; ++j) {
          k[j] = todo ? __builtin_ctzll(todo) : 0;
          if (todo) {
            todo &= todo - 1;
            nb = j + 1;
          }
        }
        float x[kLaneBatch];
#pragma unroll
        for (int j = 0; j < kLaneBatch; ++j) {
          const int uu = lane_value(u, k[j]);
          x[j] = (mine && j < nb) ? contrib[(size_t)uu * kLanes + lane] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < kLaneBatch; ++j) {
          if (j < nb) {
            const int pos = pb + k[j];
            // close the vertices whose in-edges all lie before this edge
            // (the segment's last vertex ends past every edge: r < cnt)
            while (lane_value(end, r) <= pos) {
              const int v = v0 + r;
              float nr = __fmul_rn(0.85f, acc);
              if (mine && my_seed == v) nr = __fadd_rn(nr, 0.15f);
              if (mine) rank[(size_t)v * kLanes + lane] = nr;
              const uint64_t nzm = __ballot(mine && nr != 0.0f);
              if (lane == 0) nz_out[v] = (lane_value64(nzr, r) & ~active) | nzm;
              acc = 0.0f;
              ++r;
            }
            if (mine) acc = __fadd_rn(acc, x[j]);
          }
        }
      }
    }
    for (; r < cnt; ++r) {
      const int v = v0 + r;
      float nr = __fmul_rn(0.85f, acc);
      if (mine && my_seed == v) nr = __fadd_rn(nr, 0.15f);
      if (mine) rank[(size_t)v * kLanes + lane] = nr;
      const uint64_t nzm = __ballot(mine && nr != 0.0f);
      if (lane == 0) nz_out[v] = (lane_value64(nzr, r) & ~active) | nzm;
      acc = 0.0f;
    }
  }
}

// PPR on heavy vertices: per-chunk partial sums, then their sum in chunk order.
__global__ __launch_bounds__(kSrvNT) void k_srv_ppr_heavy_partial(
    const LaneChunk* __restrict__ chunks, int n_chunks, uint64_t active,
    const int32_t* __restrict__ srcs, const float* __restrict__ contrib,
    const uint64_t* __restrict__ nz_in, float* __restrict__ partial) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (int)threadIdx.x / kWave);
  for (int c = w0; c < n_chunks; c += n_waves) {
    const int b = __builtin_amdgcn_readfirstlane(chunks[c].b);
    const int e = __builtin_amdgcn_readfirstlane(chunks[c].e);
    partial[(size_t)c * kLanes + lane] = lane_sum_in_edges(b, e, active, srcs, contrib, nz_in);
  }
}

__global__ __launch_bounds__(kSrvNT) void k_srv_ppr_heavy_combine(
    const LaneHeavy* __restrict__ heavy, int n_heavy, uint64_t active, LaneSeeds seeds,
    const float* __restrict__ partial, const uint64_t* __restrict__ nz_in,
    uint64_t* __restrict__ nz_out, float* __restrict__ rank) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  const bool mine = (active >> lane) & 1ull;
  for (int h = blockIdx.x * wpb + threadIdx.x / kWave; h < n_heavy; h += n_waves) {
    const LaneHeavy hv = heavy[h];
    float acc = 0.0f;
    for (int c = 0; c < hv.count; ++c) acc = __fadd_rn(acc, partial[(size_t)(hv.first + c) * kLanes + lane]);
    float nr = __fmul_rn(0.85f, acc);
    if (mine && seeds.vertex[lane] == hv.v) nr = __fadd_rn(nr, 0.15f);
    if (mine) rank[(size_t)hv.v * kLanes + lane] = nr;
    const uint64_t nzm = __ballot(mine && nr != 0.0f);
    if (lane == 0) nz_out[hv.v] = (nz_in[hv.v] & ~active) | nzm;
  }
}

// Checksums of the lanes in `mask` of a vertex-major [V][64] array: one wave
// per vertex reads the 64 values of a vertex at once.
template <class T>
__global__ __launch_bounds__(kSrvNT) void k_srv_checksum_lanes(int nV, uint64_t mask,
                                                               const T* __restrict__ data,
                                                               const int32_t* __restrict__ perm,
                                                               SrvCounters* __restrict__ ctr) {
  // the workgroup's waves are summed per lane in LDS first: one atomic per
  // lane and workgroup (per-wave atomics on the same 64 words serialise)
  __shared__ unsigned long long s_h[kSrvNT / kWave][kLanes];
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int n_waves = gridDim.x * wpb;
  const bool mine = (mask >> lane) & 1ull;
  unsigned long long h = 0;
  for (int v = blockIdx.x * wpb + threadIdx.x / kWave; v < nV; v += n_waves) {
    if (!mine) continue;
    const T x = data[(size_t)v * kLanes + lane];
    uint32_t bits;
    memcpy(&bits, &x, 4);
    const uint32_t id = perm ? (uint32_t)perm[v] : (uint32_t)v;  // the caller's vertex id
    h += mix64(((uint64_t)id << 32) | bits);
  }
  s_h[threadIdx.x / kWave][lane] = h;
  __syncthreads();
  if (threadIdx.x < kWave) {
    unsigned long long t = 0;
    for (int w = 0; w < wpb; ++w) t += s_h[w][lane];
    if (mine && t) atomicAdd(&ctr->checksum[lane], t);
  }
}

// Order-independent checksum of one lane's result row: sum of mix64(v, value)
// over the caller's vertex ids v (perm maps layout ids to them).
template <class T>
__global__ void k_srv_checksum(int nV, int lane_count, uint64_t mask, const T* __restrict__ data,
                               int lane_major, const int32_t* __restrict__ perm,
                               SrvCounters* __restrict__ ctr) {
  // lane_major: data[s * nV + v]; otherwise data[v * 64 + s]
  for (int s = 0; s < lane_count; ++s) {
    if (!((mask >> s) & 1ull)) continue;
    unsigned long long h = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
      const T x = lane_major ? data[(size_t)s * nV + v] : data[(size_t)v * kLanes + s];
      uint32_t bits;
      memcpy(&bits, &x, 4);
      const uint32_t id = perm ? (uint32_t)perm[v] : (uint32_t)v;
      h += mix64(((uint64_t)id << 32) | bits);
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) h += __shfl_xor(h, off, kWave);
    if (lane_id() == 0 && h) atomicAdd(&ctr->checksum[s], h);
  }
}

// Result row of lane s in the caller's ids: out[u] = value at layout id inv[u].
template <class T>
__global__ void k_srv_extract(int nV, int s, const T* __restrict__ data, int lane_major,
                              const int32_t* __restrict__ inv, T* __restrict__ out) {
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < nV; u += gridDim.x * blockDim.x) {
    const size_t v = inv ? (size_t)inv[u] : (size_t)u;
    out[u] = lane_major ? data[(size_t)s * nV + v] : data[v * kLanes + s];
  }
}

__global__ void k_permute_w(const int32_t* __restrict__ in, const int32_t* __restrict__ eid,
                            int64_t n, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[eid ? eid[i] : i];
}

// -------------------------------------------------------------------- host

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id;
  int kind;
  int32_t vertex;
  bool keep;
  Clock::time_point submitted, finished;
  bool done = false;
  uint64_t checksum = 0;
  int32_t steps = 0;
  std::vector<int32_t> result;  // raw 4-byte values (int32 or float bits)
};

struct Pool {
  int kind = 0;
  hipStream_t stream = nullptr;
  int capacity = 0;     // lanes usable under the memory budget
  uint64_t active = 0;  // lanes in use
  LaneSeeds seeds{};
  int64_t lane_req[kLanes];
  int32_t lane_steps[kLanes];
  int step = 0;         // pool-wide step counter
  size_t lane_bytes = 0;
  // device state
  uint64_t *m0 = nullptr, *m1 = nullptr, *m2 = nullptr;  // BFS visit/frontier/next; SSSP chg a/b; PPR nz a/b
  int32_t* i32 = nullptr;  // BFS depth [64][V] or SSSP dist [V][64]
  float *f0 = nullptr, *f1 = nullptr;  // PPR rank, contrib [V][64]
  SrvCounters* ctr = nullptr;
  SrvCounters* h_ctr = nullptr;  // pinned
  int32_t* h_row = nullptr;      // pinned V ints for result copies
  int parity = 0;                // SSSP/PPR ping-pong
  uint64_t last_fe = 0;          // BFS frontier / SSSP changed out-edges of the last step
};

}  // namespace gr

struct gr_server {
  gr_graph* g = nullptr;              // degree-ordered copy the kernels run on
  const int32_t* perm = nullptr;      // its ids -> the caller's ids (device)
  const int32_t* inv = nullptr;       // the caller's ids -> its ids (device)
  std::vector<int32_t> h_inv;         // host copy of inv (query seeds)
  gr_server_config cfg{};
  int32_t* w_csc = nullptr;
  uint8_t* sssp_dirty = nullptr;   // [V] flags of filtered SSSP steps (all zero between steps)
  int32_t* sssp_list = nullptr;    // [V] their light vertices
  gr::Pool pools[3];
  gr::LaneTables lanes;       // heavy-vertex chunks of the CSC
  gr::LaneTables out_lanes;   // ... of the CSR (BFS push)
  float* ppr_partial = nullptr;  // [chunks][64]
  // PPR steps: vertices below ppr_split (a multiple of 64; the degree order's
  // head) one wave each, the segments from there on one wave per 64 vertices
  int32_t ppr_split = 0;
  int32_t lane_split = 0;  // the SSSP relax walk's (graph_lane_split)
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> queue[3];
  std::unordered_map<int64_t, gr::Request> reqs;
  int64_t next_id = 1;
  std::thread workers[3];
  std::atomic<bool> stop{false};
  std::atomic<int> error{GR_OK};
  std::string error_msg;
  size_t bytes_in_use = 0, peak_bytes = 0;
  int max_lanes_in_use = 0;
  std::atomic<int64_t> steps_run{0};
  int64_t completed = 0;
  std::vector<double> latencies_ms;
  gr::Clock::time_point first_submit, last_finish;
  bool any_submit = false;
};

namespace gr {
namespace {

int alloc_dev(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_errorf(GR_ERR_NOMEM, "server: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return GR_OK;
}

unsigned grid_v(int nV) {
  const int wpb = kSrvNT / kWave;
  int g = (nV + wpb - 1) / wpb;
  return (unsigned)std::max(1, std::min(g, 8192));
}

int pool_init(gr_server* s, Pool& p, int kind) {
  const int nV = s->g->nV;
  p.kind = kind;
  GR_HIP(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
  std::fill(p.lane_req, p.lane_req + kLanes, (int64_t)-1);
  std::fill(p.lane_steps, p.lane_steps + kLanes, 0);
  GR_TRY(alloc_dev((void**)&p.ctr, sizeof(SrvCounters)));
  GR_HIP(hipHostMalloc((void**)&p.h_ctr, sizeof(SrvCounters), hipHostMallocDefault));
  GR_HIP(hipHostMalloc((void**)&p.h_row, sizeof(int32_t) * (size_t)nV, hipHostMallocDefault));
  if (kind == GR_QUERY_BFS) {
    GR_TRY(alloc_dev((void**)&p.m0, 8 * (size_t)nV));
    GR_TRY(alloc_dev((void**)&p.m1, 8 * (size_t)nV));
    GR_TRY(alloc_dev((void**)&p.m2, 8 * (size_t)nV));
    GR_TRY(alloc_dev((void**)&p.i32, 4 * (size_t)nV * kLanes));
    GR_HIP(hipMemsetAsync(p.m0, 0, 8 * (size_t)nV, p.stream));
    GR_HIP(hipMemsetAsync(p.m1, 0, 8 * (size_t)nV, p.stream));
    GR_HIP(hipMemsetAsync(p.m2, 0, 8 * (size_t)nV, p.stream));
    p.lane_bytes = 4 * (size_t)nV + 24 * (size_t)nV / kLanes;
  } else if (kind == GR_QUERY_SSSP) {
    GR_TRY(alloc_dev((void**)&p.m0, 8 * (size_t)nV));
    GR_TRY(alloc_dev((void**)&p.m1, 8 * (size_t)nV));
    GR_TRY(alloc_dev((void**)&p.i32, 4 * (size_t)nV * kLanes));
    GR_HIP(hipMemsetAsync(p.m0, 0, 8 * (size_t)nV, p.stream));
    GR_HIP(hipMemsetAsync(p.m1, 0, 8 * (size_t)nV, p.stream));
    p.lane_bytes = 4 * (size_t)nV + 16 * (size_t)nV / kLanes;
  } else {
    GR_TRY(alloc_dev((void**)&p.m0, 8 * (size_t)nV));
    GR_TRY(alloc_dev((void**)&p.m1, 8 * (size_t)nV));
    GR_TRY(alloc_dev((void**)&p.f0, 4 * (size_t)nV * kLanes));
    GR_TRY(alloc_dev((void**)&p.f1, 4 * (size_t)nV * kLanes));
    GR_HIP(hipMemsetAsync(p.m0, 0, 8 * (size_t)nV, p.stream));
    GR_HIP(hipMemsetAsync(p.m1, 0, 8 * (size_t)nV, p.stream));
    p.lane_bytes = 8 * (size_t)nV + 16 * (size_t)nV / kLanes;
  }
  const int lanes = std::max(0, std::min(kLanes, s->cfg.max_lanes_per_kind > 0
                                                     ? s->cfg.max_lanes_per_kind
                                                     : kLanes));
  p.capacity = lanes;
  GR_HIP(hipStreamSynchronize(p.stream));
  return GR_OK;
}

void pool_free(Pool& p) {
  void* ds[] = {p.m0, p.m1, p.m2, p.i32, p.f0, p.f1, p.ctr};
  for (void* d : ds)
    if (d) (void)hipFree(d);
  if (p.h_ctr) (void)hipHostFree(p.h_ctr);
  if (p.h_row) (void)hipHostFree(p.h_row);
  if (p.stream) (void)hipStreamDestroy(p.stream);
  p = Pool();
}

int free_lane(const Pool& p) {
  for (int s = 0; s < p.capacity; ++s)
    if (!((p.active >> s) & 1ull)) return s;
  return -1;
}

// Admission: under the lock, move queued requests into free lanes.
uint64_t admit(gr_server* sv, Pool& p) {
  uint64_t mask = 0;
  std::deque<int64_t>& q = sv->queue[p.kind - 1];
  while (!q.empty()) {
    const int s = free_lane(p);
    if (s < 0) break;
    if (sv->cfg.memory_budget_bytes &&
        sv->bytes_in_use + p.lane_bytes > sv->cfg.memory_budget_bytes)
      break;
    const int64_t id = q.front();
    q.pop_front();
    Request& r = sv->reqs[id];
    p.active |= 1ull << s;
    p.lane_req[s] = id;
    p.lane_steps[s] = 0;
    p.seeds.vertex[s] = sv->h_inv[r.vertex];
    p.seeds.start[s] = p.step;
    sv->bytes_in_use += p.lane_bytes;
    sv->peak_bytes = std::max(sv->peak_bytes, sv->bytes_in_use);
    mask |= 1ull << s;
  }
  return mask;
}

int launch_admit(gr_server* sv, Pool& p, uint64_t mask) {
  if (!mask) return GR_OK;
  const int nV = sv->g->nV;
  if (p.kind == GR_QUERY_BFS) {
    k_srv_bfs_clear<<<grid_1d(nV, 256), 256, 0, p.stream>>>(nV, ~mask, p.m0, p.m1, p.m2);
    k_srv_bfs_reset_depth<<<grid_v(nV), kSrvNT, 0, p.stream>>>(nV, mask, p.i32);
    k_srv_bfs_seed<<<1, kLanes, 0, p.stream>>>(mask, p.seeds, nV, p.m0, p.m1, p.i32);
  } else if (p.kind == GR_QUERY_SSSP) {
    k_srv_sssp_reset<<<grid_v(nV), kSrvNT, 0, p.stream>>>(nV, mask, 2147483647 - 100000, p.i32,
                                                          p.m0, p.m1);
    k_srv_sssp_seed<<<1, kLanes, 0, p.stream>>>(mask, p.seeds, p.i32, p.parity ? p.m1 : p.m0);
  } else {
    k_srv_ppr_reset<<<grid_v(nV), kSrvNT, 0, p.stream>>>(nV, mask, p.seeds, p.f0,
                                                         p.parity ? p.m1 : p.m0);
  }
  GR_HIP(hipGetLastError());
  return GR_OK;
}

// Enqueues one step of the pool and the read-back of its counters.
int launch_step(gr_server* sv, Pool& p) {
  gr_graph* g = sv->g;
  const int nV = g->nV;
  GR_HIP(hipMemsetAsync(p.ctr, 0, sizeof(SrvCounters), p.stream));
  if (p.kind == GR_QUERY_BFS) {
    const long long thr = (long long)((double)g->nE / 20.0);
    const int pull = (long long)p.last_fe > thr;
    k_srv_bfs_expand<<<grid_1d(nV, kSrvNT), kSrvNT, 0, p.stream>>>(
        nV, pull, p.active, g->off_out, g->dsts, g->off_in, g->srcs, p.m0, p.m1, p.m2);
    if (pull && sv->lanes.n_chunks)
      k_srv_bfs_pull_chunks<<<grid_1d(sv->lanes.n_chunks, kSrvNT / kWave), kSrvNT, 0, p.stream>>>(
          sv->lanes.chunks, sv->lanes.n_chunks, p.active, g->srcs, p.m0, p.m1, p.m2);
    if (!pull && sv->out_lanes.n_chunks)
      k_srv_bfs_push_chunks<<<grid_1d(sv->out_lanes.n_chunks, kSrvNT / kWave), kSrvNT, 0,
                              p.stream>>>(sv->out_lanes.chunks, sv->out_lanes.n_chunks, g->dsts,
                                          p.m0, p.m1, p.m2);
    k_srv_bfs_update<<<std::min<unsigned>(grid_1d(nV, kSrvNT), kAtomicGrid), kSrvNT, 0, p.stream>>>(
        nV, p.step, p.active, p.seeds, g->off_out, p.m0, p.m1, p.m2, p.i32, p.ctr);
  } else if (p.kind == GR_QUERY_SSSP) {
    const uint64_t* cin = p.parity ? p.m1 : p.m0;
    uint64_t* cout = p.parity ? p.m0 : p.m1;
    unsigned long long* any = &p.ctr->frontier_or;
    // a step visits only the out-neighbours of the changed (and newly seeded)
    // vertices when the last step's changes had few out-edges
    LaneRound rr;
    rr.prev_count = nullptr;
    rr.prev_edges = nullptr;
    rr.count = &p.ctr->changed;
    rr.edges = &p.ctr->frontier_edges;
    rr.dirty = sv->sssp_dirty;
    rr.list = sv->sssp_list;
    rr.list_len = &p.ctr->list_len;
    rr.filter_max = g->nE / 8;
    rr.host_filter = (long long)p.last_fe < rr.filter_max ? 1 : 0;
    const unsigned vgrid = std::min<unsigned>(grid_1d((int64_t)nV, kLaneNT), kAtomicGrid);
    k_lane_sssp_mark<<<vgrid, kLaneNT, 0, p.stream>>>(nV, g->off_out, g->dsts,
                                                      sv->out_lanes.chunks,
                                                      sv->out_lanes.n_chunks, cin, rr);
    k_lane_sssp_filter<<<vgrid, kLaneNT, 0, p.stream>>>(nV, g->off_in, cout, rr);
    k_lane_sssp_light<<<grid_v(nV), kLaneNT, 0, p.stream>>>(nV, g->off_in, g->off_out, g->srcs,
                                                            sv->w_csc, p.i32, cin, cout, any, rr,
                                                            sv->lane_split);
    if (sv->lanes.n_chunks) {
      k_lane_sssp_heavy<<<grid_1d(sv->lanes.n_chunks, kLaneNT / kWave), kLaneNT, 0, p.stream>>>(
          sv->lanes.chunks, sv->lanes.n_chunks, g->off_out, g->srcs, sv->w_csc, p.i32, cin, cout,
          any, rr);
      k_lane_sssp_clear_heavy<<<grid_1d(sv->lanes.n_heavy, 256), 256, 0, p.stream>>>(
          sv->lanes.heavy, sv->lanes.n_heavy, rr);
    }
    p.parity ^= 1;
  } else {
    const uint64_t* nzi = p.parity ? p.m1 : p.m0;
    uint64_t* nzo = p.parity ? p.m0 : p.m1;
    k_srv_ppr_contrib<<<grid_v(nV), kSrvNT, 0, p.stream>>>(nV, p.active, g->off_out, p.f0, p.f1);
    const int split = sv->ppr_split;
    if (split > 0)
      k_srv_ppr_step_rows<<<grid_v(split), kSrvNT, 0, p.stream>>>(0, split, p.active, p.seeds,
                                                                  g->off_in, g->srcs, p.f1, nzi,
                                                                  nzo, p.f0);
    if (split < nV)  // one wave per 64-vertex segment
      k_srv_ppr_step<<<grid_v((nV - split + kLanes - 1) / kLanes), kSrvNT, 0, p.stream>>>(
          split, nV, p.active, p.seeds, g->off_in, g->srcs, p.f1, nzi, nzo, p.f0);
    if (sv->lanes.n_chunks) {
      k_srv_ppr_heavy_partial<<<grid_1d(sv->lanes.n_chunks, kSrvNT / kWave), kSrvNT, 0, p.stream>>>(
          sv->lanes.chunks, sv->lanes.n_chunks, p.active, g->srcs, p.f1, nzi, sv->ppr_partial);
      k_srv_ppr_heavy_combine<<<grid_1d(sv->lanes.n_heavy, kSrvNT / kWave), kSrvNT, 0, p.stream>>>(
          sv->lanes.heavy, sv->lanes.n_heavy, p.active, p.seeds, sv->ppr_partial, nzi, nzo, p.f0);
    }
    p.parity ^= 1;
  }
  GR_HIP(hipGetLastError());
  GR_HIP(hipMemcpyAsync(p.h_ctr, p.ctr, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        p.stream));
  return GR_OK;
}

// After the step's stream completed: lanes that finished.
uint64_t finished_lanes(gr_server* sv, Pool& p) {
  ++p.step;
  uint64_t done = 0;
  for (int s = 0; s < kLanes; ++s)
    if ((p.active >> s) & 1ull) ++p.lane_steps[s];
  if (p.kind == GR_QUERY_BFS) {
    p.last_fe = p.h_ctr->frontier_edges;
    done = p.active & ~p.h_ctr->frontier_or;
  } else if (p.kind == GR_QUERY_SSSP) {
    p.last_fe = p.h_ctr->frontier_edges;
    done = p.active & ~p.h_ctr->frontier_or;
  } else {
    const int iters = sv->cfg.ppr_iterations > 0 ? sv->cfg.ppr_iterations : 10;
    for (int s = 0; s < kLanes; ++s)
      if (((p.active >> s) & 1ull) && p.lane_steps[s] >= iters) done |= 1ull << s;
  }
  return done;
}

int retire(gr_server* sv, Pool& p, uint64_t done) {
  if (!done) return GR_OK;
  const int nV = sv->g->nV;
  GR_HIP(hipMemsetAsync(p.ctr, 0, sizeof(SrvCounters), p.stream));
  if (p.kind == GR_QUERY_BFS || p.kind == GR_QUERY_SSSP)
    k_srv_checksum_lanes<int32_t><<<std::min(grid_v(nV), (unsigned)kAtomicGrid), kSrvNT, 0, p.stream>>>(nV, done, p.i32, sv->perm,
                                                                        p.ctr);
  else
    k_srv_checksum_lanes<float><<<std::min(grid_v(nV), (unsigned)kAtomicGrid), kSrvNT, 0, p.stream>>>(nV, done, p.f0, sv->perm,
                                                                      p.ctr);
  GR_HIP(hipGetLastError());
  GR_HIP(hipMemcpyAsync(p.h_ctr, p.ctr, sizeof(SrvCounters), hipMemcpyDeviceToHost, p.stream));
  GR_HIP(hipStreamSynchronize(p.stream));
  std::vector<int> lanes;
  for (int s = 0; s < kLanes; ++s)
    if ((done >> s) & 1ull) lanes.push_back(s);
  // result rows of the requests that keep them: extracted into a scratch row
  for (int s : lanes) {
    std::vector<int32_t> row;
    bool keep;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      keep = sv->reqs[p.lane_req[s]].keep;
    }
    if (keep) {
      int32_t* scratch = nullptr;
      GR_TRY(alloc_dev((void**)&scratch, 4 * (size_t)nV));
      if (p.kind == GR_QUERY_BFS || p.kind == GR_QUERY_SSSP)
        k_srv_extract<int32_t><<<1024, 256, 0, p.stream>>>(nV, s, p.i32, 0, sv->inv, scratch);
      else
        k_srv_extract<float><<<1024, 256, 0, p.stream>>>(nV, s, p.f0, 0, sv->inv,
                                                         (float*)scratch);
      GR_HIP(hipMemcpyAsync(p.h_row, scratch, 4 * (size_t)nV, hipMemcpyDeviceToHost, p.stream));
      GR_HIP(hipStreamSynchronize(p.stream));
      (void)hipFree(scratch);
      row.assign(p.h_row, p.h_row + nV);
    }
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      Request& r = sv->reqs[p.lane_req[s]];
      r.checksum = p.h_ctr->checksum[s];
      r.steps = p.lane_steps[s];
      r.result.swap(row);
      r.finished = now;
      r.done = true;
      sv->latencies_ms.push_back(
          std::chrono::duration<double, std::milli>(r.finished - r.submitted).count());
      sv->completed += 1;
      sv->last_finish = now;
      sv->bytes_in_use -= p.lane_bytes;
      p.active &= ~(1ull << s);  // read by the other pools' threads under the lock
      p.lane_req[s] = -1;
    }
  }
  sv->cv.notify_all();
  return GR_OK;
}

// One host thread per pool: each kind advances at its own step rate, on its
// own stream, so a slow PPR iteration does not hold BFS levels back.
void worker_loop(gr_server* sv, int k) {
  (void)hipSetDevice(0);
  Pool& p = sv->pools[k];
  while (!sv->stop.load()) {
    uint64_t adm = 0;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      adm = admit(sv, p);
      if (!p.active) {
        sv->cv.wait_for(lk, std::chrono::milliseconds(2));
        continue;
      }
      int used = 0;
      for (int j = 0; j < 3; ++j) used += __builtin_popcountll(sv->pools[j].active);
      sv->max_lanes_in_use = std::max(sv->max_lanes_in_use, used);
    }
    int rc = launch_admit(sv, p, adm);
    if (rc == GR_OK) rc = launch_step(sv, p);
    if (rc == GR_OK && hipStreamSynchronize(p.stream) != hipSuccess)
      rc = set_error(GR_ERR_HIP, "server: step failed");
    if (rc == GR_OK) rc = retire(sv, p, finished_lanes(sv, p));
    sv->steps_run += 1;
    if (rc != GR_OK) {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->error = rc;
      sv->error_msg = gr_last_error();
      sv->stop = true;
      sv->cv.notify_all();
      return;
    }
  }
}

}  // namespace
}  // namespace gr

using namespace gr;

extern "C" int gr_server_create(gr_graph* g, const int32_t* weights_device,
                                const gr_server_config* config, gr_server** out) {
  if (!g || !out) return set_error(GR_ERR_INVALID, "gr_server_create: bad arguments");
  gr_server* sv = new gr_server();
  if (config) sv->cfg = *config;
  // queries run on the degree-ordered copy of the graph (shared with engines)
  gr_graph* user_g = g;
  int rc = graph_ordered(user_g, &g);
  if (rc != GR_OK) {
    delete sv;
    return rc;
  }
  sv->g = g;
  sv->perm = user_g->perm;
  sv->inv = user_g->inv;
  sv->h_inv.resize((size_t)g->nV);
  if (g->nV && hipMemcpy(sv->h_inv.data(), sv->inv, sizeof(int32_t) * (size_t)g->nV,
                         hipMemcpyDeviceToHost) != hipSuccess)
    rc = set_error(GR_ERR_HIP, "copy of the vertex order");
  if (rc == GR_OK && weights_device && g->nE) {
    rc = alloc_dev((void**)&sv->w_csc, 4 * (size_t)g->nE);
    if (rc == GR_OK) {
      k_permute_w<<<grid_1d(g->nE, 256), 256>>>(weights_device, g->csc_eid, g->nE, sv->w_csc);
      if (hipDeviceSynchronize() != hipSuccess) rc = set_error(GR_ERR_HIP, "weights permute");
    }
  }
  if (rc == GR_OK && sv->w_csc) {
    rc = alloc_dev((void**)&sv->sssp_dirty, (size_t)g->nV + 1);
    if (rc == GR_OK) rc = alloc_dev((void**)&sv->sssp_list, 4 * ((size_t)g->nV + 1));
    if (rc == GR_OK && hipMemset(sv->sssp_dirty, 0, (size_t)g->nV + 1) != hipSuccess)
      rc = set_error(GR_ERR_HIP, "server: clearing the SSSP flags");
  }
  if (rc == GR_OK) rc = sv->lanes.build(g->h_off_in);
  if (rc == GR_OK) sv->lane_split = graph_lane_split(g);
  if (rc == GR_OK) {
    // the last 64-vertex segment from which on every segment has at most
    // kPprSegEdges in-edges and no heavy row: the segment walk's share
    const std::vector<int32_t>& off = g->h_off_in;
    int32_t split = (g->nV + kLanes - 1) / kLanes * kLanes;
    while (split > 0) {
      const int32_t a = split - kLanes, b = std::min<int32_t>(split, g->nV);
      bool light = off[b] - off[a] <= kPprSegEdges;
      for (int32_t v = a; light && v < b; ++v) light = off[v + 1] - off[v] <= kLaneHeavy;
      if (!light) break;
      split = a;
    }
    sv->ppr_split = std::min<int32_t>(split, g->nV);
  }
  if (rc == GR_OK) rc = sv->out_lanes.build(g->h_off_out);
  if (rc == GR_OK)
    rc = alloc_dev((void**)&sv->ppr_partial, sizeof(float) * kLanes * ((size_t)sv->lanes.n_chunks + 1));
  for (int k = 0; k < 3 && rc == GR_OK; ++k) rc = pool_init(sv, sv->pools[k], k + 1);
  if (rc != GR_OK) {
    for (auto& p : sv->pools) pool_free(p);
    if (sv->w_csc) (void)hipFree(sv->w_csc);
    if (sv->sssp_dirty) (void)hipFree(sv->sssp_dirty);
    if (sv->sssp_list) (void)hipFree(sv->sssp_list);
    if (sv->ppr_partial) (void)hipFree(sv->ppr_partial);
    delete sv;
    return rc;
  }
  for (int k = 0; k < 3; ++k) sv->workers[k] = std::thread(worker_loop, sv, k);
  *out = sv;
  return GR_OK;
}

extern "C" int gr_server_destroy(gr_server* sv) {
  if (!sv) return GR_OK;
  sv->stop = true;
  sv->cv.notify_all();
  for (auto& w : sv->workers)
    if (w.joinable()) w.join();
  for (auto& p : sv->pools) pool_free(p);
  if (sv->w_csc) (void)hipFree(sv->w_csc);
  if (sv->sssp_dirty) (void)hipFree(sv->sssp_dirty);
  if (sv->sssp_list) (void)hipFree(sv->sssp_list);
  if (sv->ppr_partial) (void)hipFree(sv->ppr_partial);
  delete sv;
  return GR_OK;
}

extern "C" int gr_server_submit(gr_server* sv, int kind, int32_t vertex, int keep_result,
                                int64_t* request_id) {
  if (!sv || !request_id) return set_error(GR_ERR_INVALID, "gr_server_submit: bad arguments");
  if (kind < GR_QUERY_BFS || kind > GR_QUERY_PPR)
    return set_errorf(GR_ERR_INVALID, "unknown query kind %d", kind);
  if (vertex < 0 || vertex >= sv->g->nV)
    return set_errorf(GR_ERR_INVALID, "query vertex %d out of range", vertex);
  if (kind == GR_QUERY_SSSP && !sv->w_csc)
    return set_error(GR_ERR_INVALID, "server has no edge weights for SSSP");
  if (sv->error.load() != GR_OK) return set_error(sv->error.load(), "server stopped after an error");
  const Pool& p = sv->pools[kind - 1];
  if (p.capacity == 0 ||
      (sv->cfg.memory_budget_bytes && p.lane_bytes > sv->cfg.memory_budget_bytes))
    return set_errorf(GR_ERR_BUDGET, "query needs %zu bytes, over the server's budget",
                      p.lane_bytes);
  std::lock_guard<std::mutex> lk(sv->mu);
  Request r;
  r.id = sv->next_id++;
  r.kind = kind;
  r.vertex = vertex;
  r.keep = keep_result != 0;
  r.submitted = Clock::now();
  if (!sv->any_submit) {
    sv->first_submit = r.submitted;
    sv->any_submit = true;
  }
  *request_id = r.id;
  sv->queue[kind - 1].push_back(r.id);
  sv->reqs.emplace(r.id, std::move(r));
  sv->cv.notify_all();
  return GR_OK;
}

extern "C" int gr_server_wait(gr_server* sv, int64_t request_id, void* result, uint64_t* checksum,
                              double* latency_ms, int32_t* steps) {
  if (!sv) return set_error(GR_ERR_INVALID, "server is NULL");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(request_id);
  if (it == sv->reqs.end()) return set_errorf(GR_ERR_INVALID, "unknown request %lld", (long long)request_id);
  sv->cv.wait(lk, [&] { return it->second.done || sv->error.load() != GR_OK; });
  if (!it->second.done) return set_errorf(sv->error.load(), "server error: %s", sv->error_msg.c_str());
  Request& r = it->second;
  if (result && !r.result.empty()) ::memcpy(result, r.result.data(), 4 * r.result.size());
  if (result && r.result.empty()) return set_error(GR_ERR_STATE, "request did not keep its result");
  if (checksum) *checksum = r.checksum;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.finished - r.submitted).count();
  if (steps) *steps = r.steps;
  r.result.clear();
  r.result.shrink_to_fit();
  return GR_OK;
}

extern "C" int gr_server_get_stats(gr_server* sv, gr_server_stats* out) {
  if (!sv || !out) return set_error(GR_ERR_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(sv->mu);
  gr_server_stats st{};
  st.completed = sv->completed;
  st.submitted = sv->next_id - 1;
  st.steps = sv->steps_run.load();
  st.max_lanes_in_use = sv->max_lanes_in_use;
  st.peak_bytes = (int64_t)sv->peak_bytes;
  if (!sv->latencies_ms.empty()) {
    std::vector<double> l = sv->latencies_ms;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      size_t i = (size_t)std::min<double>(l.size() - 1, std::ceil(q * l.size()) - 1);
      return l[i];
    };
    st.p50_ms = pct(0.50);
    st.p99_ms = pct(0.99);
    st.max_ms = l.back();
    if (sv->any_submit) {
      const double span = std::chrono::duration<double>(sv->last_finish - sv->first_submit).count();
      st.throughput_rps = span > 0 ? sv->completed / span : 0.0;
    }
  }
  *out = st;
  return GR_OK;
}

extern "C" int gr_server_load_test(gr_server* sv, double rate_per_s, int n_requests,
                                   const int32_t* vertices, const int32_t* kinds, int keep_results,
                                   int64_t* request_ids) {
  if (!sv || rate_per_s <= 0 || n_requests < 0 || !vertices || !kinds)
    return set_error(GR_ERR_INVALID, "gr_server_load_test: bad arguments");
  std::mt19937_64 rng(12345);
  std::exponential_distribution<double> gap(rate_per_s);
  auto t = Clock::now();
  for (int i = 0; i < n_requests; ++i) {
    t += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(gap(rng)));
    std::this_thread::sleep_until(t);
    int64_t id = 0;
    GR_TRY(gr_server_submit(sv, kinds[i], vertices[i], keep_results, &id));
    if (request_ids) request_ids[i] = id;
  }
  return GR_OK;
}
