// Batched multi-source traversals: up to 64 BFS or SSSP sources share one
// pass over the graph, one bit (BFS) or one lane (SSSP) per source.
//
// Multi-source BFS: per vertex three 64-bit masks live in HBM: visit
// (sources that reached it), frontier (sources for which it is on the
// current level) and next.  Each level is an expand phase and an update
// phase:
//   expand, push  every frontier vertex ORs its frontier mask into next[u] of
//                 its CSR out-neighbours (skipping bits next[u] or visit[u]
//                 already hold);
//   expand, pull  every vertex with unvisited sources ORs the frontier masks of
//                 its CSC in-neighbours, stopping once all of its unvisited
//                 sources are covered (thread per vertex; a wave per vertex
//                 above kMsHeavy in-edges);
//   update        new = next & ~visit; visit |= new; frontier = new; the
//                 level is written for every new (source, vertex) pair.
// Rows above kLaneHeavy edges are cut into chunks spread over all waves.
// The default runs one expand and one update launch per level; k_msbfs runs
// every level in ONE persistent cooperative launch with grid barriers
// between the phases (the same level bodies).  The direction is chosen per
// level from the frontier's out-edge count.  Depths equal those of
// independent plain runs (the BFS level of a vertex does not depend on the
// schedule), which the tests check.
//
// Multi-source SSSP (k_lane_sssp_light/heavy in gr_lanes.h): Bellman-Ford
// rounds over 64 lanes, one wave per destination vertex, lane s relaxing
// source s; only in-edges whose source changed in the previous round are read
// (a source that did not change cannot lower a distance it already lowered),
// so distances equal the plain engine's fixed point.  Rounds are recorded
// once into a hipGraph and replayed.
#include <hip/hip_cooperative_groups.h>

#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "gr_device.h"
#include "gr_gather.h"
#include "gr_lanes.h"

namespace cg = cooperative_groups;

namespace gr {

constexpr int kBatchNT = 256;
constexpr int kMsHeavy = 64;  // out-/in-degree above which a wave cooperates on a vertex

struct MsCounters {
  unsigned long long new_vertices[3];   // per level (ring of 3 slots)
  unsigned long long frontier_edges[3]; // out-edges of the new frontier
  int32_t levels;
  int32_t error;
};

// Grid barrier that first completes this wave's own memory operations.
// The workgroup barrier inside grid.sync() does not wait for the
// fire-and-forget atomics (counter adds, atomicOr into next) of the waves
// that are not the workgroup's barrier leader, so such an atomic could land
// after another workgroup has already read the location.  Every workgroup
// must read the same level counts (they decide when all leave the level
// loop), hence the agent-scope fence in every wave before the barrier.
__device__ __forceinline__ void grid_barrier(cg::grid_group& grid) {
  __threadfence();
  grid.sync();
}

// OR-reduction of a 64-bit value over the wave.
__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    lo |= __shfl_xor(lo, off, kWave);
    hi |= __shfl_xor(hi, off, kWave);
  }
  return ((uint64_t)hi << 32) | lo;
}

__global__ void k_ms_init(int nV, uint64_t* __restrict__ visit, uint64_t* __restrict__ frontier,
                          uint64_t* __restrict__ next, uint32_t* __restrict__ fbits) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
    visit[v] = 0;
    frontier[v] = 0;
    next[v] = 0;
    if (fbits && (v & 31) == 0) fbits[v >> 5] = 0;
  }
}

// depth_vs != 0: depth is vertex-major [V][64]; otherwise source-major [ns][V]
__global__ void k_ms_sources(int ns, const int32_t* __restrict__ sources, int nV,
                             uint64_t* __restrict__ visit, uint64_t* __restrict__ frontier,
                             uint32_t* __restrict__ fbits, int32_t* __restrict__ depth,
                             int depth_vs, const int32_t* __restrict__ off_out,
                             MsCounters* __restrict__ ctr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long fe = 0;
  for (int s = 0; s < ns; ++s) {
    const int v = sources[s];
    if (frontier[v] == 0) fe += (unsigned long long)(off_out[v + 1] - off_out[v]);
    visit[v] |= 1ull << s;
    frontier[v] |= 1ull << s;
    if (fbits) fbits[v >> 5] |= 1u << (v & 31);
    depth[depth_vs ? (size_t)v * kWave + s : (size_t)s * nV + v] = 0;
  }
  for (int i = 0; i < 3; ++i) {
    ctr->new_vertices[i] = 0;
    ctr->frontier_edges[i] = 0;
  }
  ctr->new_vertices[0] = (unsigned long long)ns;
  ctr->frontier_edges[0] = fe;
  ctr->levels = 0;
  ctr->error = 0;
}

// ------------------------------------------------------- MS-BFS level bodies
//
// One level is an expand phase (push or pull) and an update phase.  The
// bodies below are grid-strided from wave `w0` over `nw` waves, so that the
// per-level launches (k_ms_push ... k_ms_update, the host reading the level's
// two counters after the update) and the persistent launch (k_msbfs, grid
// barriers between the phases) run the same code.  fbits is the frontier
// bitmap (bit v = frontier[v] != 0), written by the update: push reads the
// 8-byte frontier mask only of frontier vertices, and pull may skip the mask
// of in-neighbours outside the frontier.  Rows above kLaneHeavy edges are
// left to the chunk bodies, which spread them over all waves in kLaneChunk
// pieces (a single wave per hub would be the critical path).

struct MsLevelCounters {
  unsigned long long new_vertices;
  unsigned long long frontier_edges;
};

__device__ __forceinline__ void ms_push_body(int w0, int nw, int nV,
                                             const int32_t* __restrict__ off_out,
                                             const int32_t* __restrict__ dsts,
                                             const uint64_t* __restrict__ visit,
                                             const uint64_t* __restrict__ frontier,
                                             const uint32_t* __restrict__ fbits,
                                             uint64_t* __restrict__ next) {
  const int lane = lane_id();
  for (int base = w0 * kWave; base < nV; base += nw * kWave) {
    const int v = base + lane;
    const uint32_t word = (v < nV) ? fbits[v >> 5] : 0u;
    const bool inf = (word >> (v & 31)) & 1u;
    if (!__ballot(inf)) continue;
    uint64_t f = inf ? frontier[v] : 0ull;
    int b = 0, e = 0;
    if (f) {
      b = off_out[v];
      e = off_out[v + 1];
      if (e - b > kLaneHeavy) f = 0;  // chunked: ms_push_chunk_body
    }
    const bool light = (e - b) <= kMsHeavy;
    if (f && light) {
      // four out-edges per step (independent loads in flight); an atomic only
      // when next[u] does not already hold the bits (hub targets get many)
      for (int p0 = b; p0 < e; p0 += 4) {
        int u[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = p0 + j < e ? dsts[p0 + j] : -1;
        uint64_t want[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) want[j] = u[j] >= 0 ? (f & ~visit[u[j]] & ~next[u[j]]) : 0ull;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (want[j]) atomicOr((unsigned long long*)&next[u[j]], (unsigned long long)want[j]);
      }
    }
    // medium rows one wave each; the row's bounds and mask are read out of
    // its lane as scalars, so the walk is a scalar loop
    uint64_t heavy = __ballot(f && !light);
    while (heavy) {
      const int hl = __builtin_ctzll(heavy);
      heavy &= heavy - 1;
      const int hb = lane_value(b, hl), he = lane_value(e, hl);
      const uint64_t hf = lane_value64(f, hl);
      for (int p = hb + lane; p < he; p += kWave) {
        const int u = dsts[p];
        const uint64_t want = hf & ~visit[u] & ~next[u];
        if (want) atomicOr((unsigned long long*)&next[u], (unsigned long long)want);
      }
    }
  }
}

__device__ __forceinline__ uint64_t ms_mask_of(int u, const uint64_t* __restrict__ frontier,
                                               const uint32_t* __restrict__ filter) {
  if (filter && !((filter[u >> 5] >> (u & 31)) & 1u)) return 0ull;
  return frontier[u];
}

// filter: the frontier bitmap, or nullptr to read every in-neighbour's mask
__device__ __forceinline__ void ms_pull_body(int w0, int nw, int nV, uint64_t all,
                                             const int32_t* __restrict__ off_in,
                                             const int32_t* __restrict__ srcs,
                                             const uint64_t* __restrict__ visit,
                                             const uint64_t* __restrict__ frontier,
                                             const uint32_t* __restrict__ filter,
                                             uint64_t* __restrict__ next) {
  const int lane = lane_id();
  for (int base = w0 * kWave; base < nV; base += nw * kWave) {
    const int v = base + lane;
    uint64_t need = v < nV ? (all & ~visit[v]) : 0ull;
    int b = 0, e = 0;
    if (need) {
      b = off_in[v];
      e = off_in[v + 1];
      if (e - b > kLaneHeavy) need = 0;  // chunked: ms_pull_chunk_body (atomicOr)
    }
    const bool light = (e - b) <= kMsHeavy;
    if (need && light) {
      uint64_t acc = 0;
      for (int p0 = b; p0 < e; p0 += 4) {  // four in-edges per step: loads in flight
        int u[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = p0 + j < e ? srcs[p0 + j] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (u[j] >= 0) acc |= ms_mask_of(u[j], frontier, filter);
        if ((acc & need) == need) break;
      }
      if (acc & need) next[v] = acc & need;
    }
    uint64_t heavy = __ballot(need && !light);
    while (heavy) {
      const int hl = __builtin_ctzll(heavy);
      heavy &= heavy - 1;
      const int hb = lane_value(b, hl), he = lane_value(e, hl);
      const uint64_t hn = lane_value64(need, hl);
      uint64_t hacc = 0;
      for (int pb = hb; pb < he; pb += kWave) {  // scalar trip count
        const int p = pb + lane;
        const uint64_t m = p < he ? ms_mask_of(srcs[p], frontier, filter) : 0ull;
        hacc |= wave_or64(m);
        if ((hacc & hn) == hn) break;
      }
      if (lane == hl && (hacc & hn)) next[v] = hacc & hn;
    }
  }
}

__device__ __forceinline__ void ms_push_chunk_body(int w0, int nw,
                                                   const LaneChunk* __restrict__ chunks,
                                                   int n_chunks, const int32_t* __restrict__ dsts,
                                                   const uint64_t* __restrict__ visit,
                                                   const uint64_t* __restrict__ frontier,
                                                   const uint32_t* __restrict__ fbits,
                                                   uint64_t* __restrict__ next) {
  const int lane = lane_id();
  for (int c = w0; c < n_chunks; c += nw) {
    const LaneChunk ch = chunks[c];
    if (!((fbits[ch.v >> 5] >> (ch.v & 31)) & 1u)) continue;
    const uint64_t f = frontier[ch.v];
    for (int p = ch.b + lane; p < ch.e; p += kWave) {
      const int u = dsts[p];
      const uint64_t want = f & ~visit[u] & ~next[u];
      if (want) atomicOr((unsigned long long*)&next[u], (unsigned long long)want);
    }
  }
}

__device__ __forceinline__ void ms_pull_chunk_body(int w0, int nw,
                                                   const LaneChunk* __restrict__ chunks,
                                                   int n_chunks, uint64_t all,
                                                   const int32_t* __restrict__ srcs,
                                                   const uint64_t* __restrict__ visit,
                                                   const uint64_t* __restrict__ frontier,
                                                   uint64_t* __restrict__ next) {
  const int lane = lane_id();
  for (int c = w0; c < n_chunks; c += nw) {
    const LaneChunk ch = chunks[c];
    const uint64_t need = all & ~visit[ch.v];
    if (!need) continue;
    uint64_t acc = 0;
    for (int pb = ch.b; pb < ch.e; pb += kWave) {  // wave-uniform trip count
      const int p = pb + lane;
      acc |= wave_or64(p < ch.e ? frontier[srcs[p]] : 0ull);
      if ((acc & need) == need) break;
    }
    if (lane == 0 && (acc & need))
      atomicOr((unsigned long long*)&next[ch.v], (unsigned long long)(acc & need));
  }
}

// The update of every vertex: new = next & ~visit; visit |= new; frontier =
// new; the level is written for every new (source, vertex) pair into the
// vertex-major depth rows; the new frontier's vertex and out-edge counts are
// added to *nv_out / *fe_out once per workgroup.  Every thread of the
// workgroup calls it (it synchronises the workgroup once).
__device__ __forceinline__ void ms_update_body(int w0, int nw, int nV, uint64_t all, int level,
                                               const int32_t* __restrict__ off_out,
                                               uint64_t* __restrict__ visit,
                                               uint64_t* __restrict__ frontier,
                                               uint64_t* __restrict__ next,
                                               uint32_t* __restrict__ fbits,
                                               int32_t* __restrict__ depth_vs,
                                               unsigned long long* __restrict__ nv_out,
                                               unsigned long long* __restrict__ fe_out) {
  __shared__ unsigned long long s_nv[kBatchNT / kWave], s_fe[kBatchNT / kWave];
  const int lane = lane_id();
  // the level counts are summed per wave over its vertices, then per
  // workgroup: one atomic pair per workgroup (not per wave) on the counters
  unsigned long long my_nv = 0, my_fe = 0;
  for (int base = w0 * kWave; base < nV; base += nw * kWave) {
    const int v = base + lane;
    uint64_t nwm = 0;
    int deg = 0;
    if (v < nV) {
      const uint64_t vis = visit[v];
      const uint64_t nxt = next[v];
      nwm = nxt & ~vis & all;
      if (nxt) next[v] = 0;
      if (nwm || frontier[v]) frontier[v] = nwm;
      if (nwm) {
        visit[v] = vis | nwm;
        deg = off_out[v + 1] - off_out[v];
      }
    }
    const uint64_t nb = __ballot(nwm != 0);
    // depths of the new (source, vertex) pairs: for each vertex of the wave
    // with new bits, lane s writes the vertex's row entry s (one masked,
    // coalesced 256-B store per vertex)
    uint64_t rows = nb;
    while (rows) {
      const int r = __builtin_ctzll(rows);
      rows &= rows - 1;
      const uint64_t m = lane_value64(nwm, r);
      if ((m >> lane) & 1ull) depth_vs[(size_t)(base + r) * kWave + lane] = level + 1;
    }
    if ((lane & 31) == 0 && v < nV) fbits[v >> 5] = (uint32_t)(nb >> lane);
    unsigned long long fe = (unsigned long long)deg;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) fe += __shfl_xor(fe, off, kWave);
    my_nv += (unsigned long long)__popcll(nb);
    my_fe += fe;
  }
  if (lane == 0) {
    s_nv[threadIdx.x / kWave] = my_nv;
    s_fe[threadIdx.x / kWave] = my_fe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long nv = 0, fe = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) {
      nv += s_nv[i];
      fe += s_fe[i];
    }
    if (nv) {
      atomicAdd(nv_out, nv);
      atomicAdd(fe_out, fe);
    }
  }
}

__device__ __forceinline__ int ms_wave0() {
  return blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
}
__device__ __forceinline__ int ms_waves() { return gridDim.x * (blockDim.x / kWave); }

__global__ __launch_bounds__(kBatchNT) void k_ms_push(int nV, const int32_t* __restrict__ off_out,
                                                      const int32_t* __restrict__ dsts,
                                                      const uint64_t* __restrict__ visit,
                                                      const uint64_t* __restrict__ frontier,
                                                      const uint32_t* __restrict__ fbits,
                                                      uint64_t* __restrict__ next) {
  ms_push_body(ms_wave0(), ms_waves(), nV, off_out, dsts, visit, frontier, fbits, next);
}

__global__ __launch_bounds__(kBatchNT) void k_ms_pull(int nV, uint64_t all,
                                                      const int32_t* __restrict__ off_in,
                                                      const int32_t* __restrict__ srcs,
                                                      const uint64_t* __restrict__ visit,
                                                      const uint64_t* __restrict__ frontier,
                                                      const uint32_t* __restrict__ filter,
                                                      uint64_t* __restrict__ next) {
  ms_pull_body(ms_wave0(), ms_waves(), nV, all, off_in, srcs, visit, frontier, filter, next);
}

__global__ __launch_bounds__(kBatchNT) void k_ms_push_chunks(
    const LaneChunk* __restrict__ chunks, int n_chunks, const int32_t* __restrict__ dsts,
    const uint64_t* __restrict__ visit, const uint64_t* __restrict__ frontier,
    const uint32_t* __restrict__ fbits, uint64_t* __restrict__ next) {
  ms_push_chunk_body(ms_wave0(), ms_waves(), chunks, n_chunks, dsts, visit, frontier, fbits, next);
}

__global__ __launch_bounds__(kBatchNT) void k_ms_pull_chunks(
    const LaneChunk* __restrict__ chunks, int n_chunks, uint64_t all,
    const int32_t* __restrict__ srcs, const uint64_t* __restrict__ visit,
    const uint64_t* __restrict__ frontier, uint64_t* __restrict__ next) {
  ms_pull_chunk_body(ms_wave0(), ms_waves(), chunks, n_chunks, all, srcs, visit, frontier, next);
}

__global__ __launch_bounds__(kBatchNT) void k_ms_update(
    int nV, uint64_t all, int level, const int32_t* __restrict__ off_out,
    uint64_t* __restrict__ visit, uint64_t* __restrict__ frontier, uint64_t* __restrict__ next,
    uint32_t* __restrict__ fbits, int32_t* __restrict__ depth_vs,
    MsLevelCounters* __restrict__ ctr) {
  ms_update_body(ms_wave0(), ms_waves(), nV, all, level, off_out, visit, frontier, next, fbits,
                 depth_vs, &ctr->new_vertices, &ctr->frontier_edges);
}

// Every level in ONE cooperative launch: the same expand and update bodies,
// separated by grid barriers.  The level's counts live in a ring of three
// slots: level L reads slot L % 3 (written by the previous update, or by
// k_ms_sources), workgroup 0 clears slot (L + 1) % 3 before the barrier that
// precedes the update which adds into it.  Every workgroup reads the counts
// with agent-scope atomic loads after a fenced grid barrier, so all of them
// take the same direction and leave the level loop at the same level.
__global__ __launch_bounds__(kBatchNT) void k_msbfs(
    int nV, long long nE, int ns, int mode, double alpha, int filter,
    const int32_t* __restrict__ off_out, const int32_t* __restrict__ dsts,
    const int32_t* __restrict__ off_in, const int32_t* __restrict__ srcs,
    uint64_t* __restrict__ visit, uint64_t* __restrict__ frontier, uint64_t* __restrict__ next,
    uint32_t* __restrict__ fbits, int32_t* __restrict__ depth_vs,
    const LaneChunk* __restrict__ in_chunks, int n_in_chunks,
    const LaneChunk* __restrict__ out_chunks, int n_out_chunks, MsCounters* __restrict__ ctr,
    int max_levels) {
  cg::grid_group grid = cg::this_grid();
  const uint64_t all = ns >= 64 ? ~0ull : ((1ull << ns) - 1);
  const int w0 = ms_wave0(), nw = ms_waves();
  const long long threshold = (long long)((double)nE / (alpha > 0 ? alpha : 1.0));
  for (int level = 0; level < max_levels; ++level) {
    const int slot = level % 3;
    const unsigned long long nv_count =
        __hip_atomic_load(&ctr->new_vertices[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long fe_count =
        __hip_atomic_load(&ctr->frontier_edges[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (nv_count == 0) break;
    const bool pull = mode == 2 || (mode == 0 && (long long)fe_count > threshold);
    const int nx = (level + 1) % 3;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      __hip_atomic_store(&ctr->new_vertices[nx], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctr->frontier_edges[nx], 0ull, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    if (pull) {
      ms_pull_body(w0, nw, nV, all, off_in, srcs, visit, frontier, filter ? fbits : nullptr,
                   next);
      ms_pull_chunk_body(w0, nw, in_chunks, n_in_chunks, all, srcs, visit, frontier, next);
    } else {
      ms_push_body(w0, nw, nV, off_out, dsts, visit, frontier, fbits, next);
      ms_push_chunk_body(w0, nw, out_chunks, n_out_chunks, dsts, visit, frontier, fbits, next);
    }
    grid_barrier(grid);
    ms_update_body(w0, nw, nV, all, level, off_out, visit, frontier, next, fbits, depth_vs,
                   &ctr->new_vertices[nx], &ctr->frontier_edges[nx]);
    grid_barrier(grid);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->levels = level + 1;
  }
}

// depth_vs ([V][64], layout ids; valid where the visit bit is set) ->
// out[s * nV + u] in the caller's ids, -1 where unreached.  A workgroup moves
// 64 consecutive caller ids through an LDS tile: their layout ids go to LDS
// first, then each of the four waves reads its 16 vertex rows of 256 B (and
// visit masks) with all loads in flight at once, then writes 64 consecutive
// depths per source.
__global__ __launch_bounds__(256) void k_ms_depth_out(int nV, int ns,
                                                      const int32_t* __restrict__ depth_vs,
                                                      const uint64_t* __restrict__ visit,
                                                      const int32_t* __restrict__ inv,
                                                      int32_t* __restrict__ out) {
  constexpr int kRows = kWave / 4;  // rows per wave
  __shared__ int32_t tile[kWave][kWave + 1];
  __shared__ int32_t s_v[kWave];
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  for (int u0 = blockIdx.x * kWave; u0 < nV; u0 += gridDim.x * kWave) {
    if (threadIdx.x < kWave) {
      const int u = u0 + threadIdx.x;
      s_v[threadIdx.x] = u < nV ? inv[u] : -1;
    }
    __syncthreads();
    int32_t d[kRows];
    uint64_t vis[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
      const int v = s_v[w + 4 * k];
      vis[k] = v >= 0 ? visit[v] : 0ull;
      d[k] = v >= 0 ? depth_vs[(size_t)v * kWave + lane] : -1;
    }
#pragma unroll
    for (int k = 0; k < kRows; ++k) tile[w + 4 * k][lane] = ((vis[k] >> lane) & 1ull) ? d[k] : -1;
    __syncthreads();
    for (int s = w; s < ns; s += 4) {
      const int u = u0 + lane;
      if (u < nV) out[(size_t)s * nV + u] = tile[lane][s];
    }
    __syncthreads();
  }
}

// Pull for the 64 vertices base .. base+63 (lane l <-> vertex base + l): the
// wave streams their contiguous CSC segment; an edge contributes
// frontier[u] when u's bit is set in the frontier bitmap; per-row ORs are
// formed with a segmented shuffle reduction and land in the row's lane.
// Rows whose unvisited sources are all covered, and heavy rows (done as
// chunks), are jumped over.  Returns next for the lane's own vertex.
__device__ __forceinline__ uint64_t ms_pull_rows(int base, int nV, uint64_t all, int heavy_deg,
                                                 const int32_t* __restrict__ off_in,
                                                 const int32_t* __restrict__ srcs,
                                                 const uint64_t* __restrict__ visit,
                                                 const uint64_t* __restrict__ frontier,
                                                 const uint32_t* __restrict__ fbits) {
  const int lane = lane_id();
  const int cnt = min(kWave, nV - base);
  const bool mine = lane < cnt;
  const uint64_t need = mine ? (all & ~visit[base + lane]) : 0ull;
  int my_off = INT_MAX, my_end = INT_MAX;
  if (mine) {
    my_off = off_in[base + lane];
    my_end = off_in[base + lane + 1];
  }
  const bool heavy = mine && (my_end - my_off) > heavy_deg;
  const uint64_t todo_rows = __ballot(need != 0 && my_end > my_off && !heavy);
  uint64_t acc = 0;
  if (!todo_rows) return 0;
  uint64_t open = todo_rows;  // rows still to be completed
  const int first = __builtin_ctzll(todo_rows);
  const int last = 63 - __builtin_clzll(todo_rows);
  int pb = __shfl(my_off, first, kWave);
  const int seg1 = __shfl(my_end, last, kWave);
  while (pb < seg1 && open) {
    const uint64_t started = __ballot(mine && my_off <= pb);
    const int r0 = 63 - __builtin_clzll(started);
    if (!((open >> r0) & 1ull)) {  // row done, heavy, empty or needless: jump
      pb = __shfl(my_end, r0, kWave);
      continue;
    }
    const int p = pb + lane;
    uint64_t val = 0;
    if (p < seg1) {
      const int u = srcs[p];
      if ((fbits[u >> 5] >> (u & 31)) & 1u) val = frontier[u];
    }
    // row of each lane's edge: last row starting at or before p
    int r = 0;
#pragma unroll
    for (int st = 32; st >= 1; st >>= 1) {
      const int c = r + st;
      const int oc = __shfl(my_off, c & 63, kWave);
      if (c < cnt && oc <= p) r = c;
    }
    // segmented OR over runs of equal rows (suffix OR within each run)
    uint32_t lo = (uint32_t)val, hi = (uint32_t)(val >> 32);
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int orow = __shfl_down(r, off, kWave);
      const uint32_t olo = __shfl_down(lo, off, kWave);
      const uint32_t ohi = __shfl_down(hi, off, kWave);
      if (lane + off < kWave && orow == r) {
        lo |= olo;
        hi |= ohi;
      }
    }
    // the head lane of row l's run is max(off_l - pb, 0)
    const int head = min(max(my_off - pb, 0), 63);
    const uint32_t hlo = __shfl(lo, head, kWave);
    const uint32_t hhi = __shfl(hi, head, kWave);
    const uint32_t hrow = __shfl(r, head, kWave);
    if (mine && my_off < pb + kWave && my_end > pb && (int)hrow == lane)
      acc |= ((uint64_t)hhi << 32) | hlo;
    const bool done_row = (acc & need) == need;
    open &= ~__ballot(done_row);
    pb += kWave;
  }
  // a heavy row may have picked up part of its edges here; its result comes
  // only from the chunk pass (atomicOr), never from a plain store
  return heavy ? 0ull : (acc & need);
}

// Streaming-pull variant (frontier bitmap, wave per 64 rows, heavy-row
// chunks).  Disabled: see kStreamPullEnabled.
__global__ __launch_bounds__(kBatchNT) void k_msbfs_stream(
    int nV, long long nE, int ns, int mode, double alpha, const int32_t* __restrict__ off_out,
    const int32_t* __restrict__ dsts, const int32_t* __restrict__ off_in,
    const int32_t* __restrict__ srcs, uint64_t* __restrict__ visit,
    uint64_t* __restrict__ frontier, uint64_t* __restrict__ next, uint32_t* __restrict__ fbits,
    int32_t* __restrict__ depth, const LaneChunk* __restrict__ chunks, int n_chunks,
    MsCounters* __restrict__ ctr, int max_levels) {
  cg::grid_group grid = cg::this_grid();
  const uint64_t all = ns >= 64 ? ~0ull : ((1ull << ns) - 1);
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int wave_id = blockIdx.x * wpb + threadIdx.x / kWave;
  const int n_waves = gridDim.x * wpb;
  const long long threshold = (long long)((double)nE / (alpha > 0 ? alpha : 1.0));
  for (int level = 0; level < max_levels; ++level) {
    const int slot = level % 3;
    // Coherent (atomic) loads after the grid barrier: every workgroup must see
    // the same counts so that all of them leave the loop at the same level
    // and pick the same direction.
    const unsigned long long nv_count =
        __hip_atomic_load(&ctr->new_vertices[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long fe_count =
        __hip_atomic_load(&ctr->frontier_edges[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (nv_count == 0) break;
    const bool pull = mode == 2 || (mode == 0 && (long long)fe_count > threshold);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const int nx = (level + 1) % 3;
      __hip_atomic_store(&ctr->new_vertices[nx], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctr->frontier_edges[nx], 0ull, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    // ---------------------------------------------------------- expand
    if (!pull) {
      for (int base = wave_id * kWave; base < nV; base += n_waves * kWave) {
        const int v = base + lane;
        const uint32_t word = (v < nV) ? fbits[v >> 5] : 0u;
        const bool inf = (word >> (v & 31)) & 1u;
        if (!__ballot(inf)) continue;
        const uint64_t f = inf ? frontier[v] : 0ull;
        int b = 0, e = 0;
        if (f) {
          b = off_out[v];
          e = off_out[v + 1];
        }
        const bool light = (e - b) <= kMsHeavy;
        if (f && light) {
          for (int p = b; p < e; ++p) {
            const int u = dsts[p];
            if (f & ~visit[u]) atomicOr((unsigned long long*)&next[u], (unsigned long long)f);
          }
        }
        uint64_t heavy = __ballot(f && !light);
        while (heavy) {
          const int hl = __builtin_ctzll(heavy);
          heavy &= heavy - 1;
          const int hb = __shfl(b, hl, kWave), he = __shfl(e, hl, kWave);
          const uint32_t flo = __shfl((uint32_t)f, hl, kWave);
          const uint32_t fhi = __shfl((uint32_t)(f >> 32), hl, kWave);
          const uint64_t hf = ((uint64_t)fhi << 32) | flo;
          for (int p = hb + lane; p < he; p += kWave) {
            const int u = dsts[p];
            if (hf & ~visit[u]) atomicOr((unsigned long long*)&next[u], (unsigned long long)hf);
          }
        }
      }
    } else {
      for (int base = wave_id * kWave; base < nV; base += n_waves * kWave) {
        const uint64_t nx = ms_pull_rows(base, nV, all, kLaneHeavy, off_in, srcs, visit,
                                         frontier, fbits);
        if (nx) next[base + lane] = nx;  // heavy rows are left to the chunk pass
      }
      // heavy rows: chunks of in-edges spread over all waves, ORed in atomically
      for (int c = wave_id; c < n_chunks; c += n_waves) {
        const LaneChunk ch = chunks[c];
        const uint64_t need = all & ~visit[ch.v];
        if (!need) continue;
        uint64_t acc = 0;
        for (int pb = ch.b; pb < ch.e; pb += kWave) {  // wave-uniform trip count
          const int p = pb + lane;
          uint64_t m = 0;
          if (p < ch.e) {
            const int u = srcs[p];
            if ((fbits[u >> 5] >> (u & 31)) & 1u) m = frontier[u];
          }
          acc |= wave_or64(m);
          if ((acc & need) == need) break;
        }
        if (lane == 0 && (acc & need))
          atomicOr((unsigned long long*)&next[ch.v], (unsigned long long)(acc & need));
      }
    }
    grid_barrier(grid);
    // ---------------------------------------------------------- update
    const int slot_next = (level + 1) % 3;
    for (int base = wave_id * kWave; base < nV; base += n_waves * kWave) {
      const int v = base + lane;
      uint64_t nw = 0;
      int deg = 0;
      if (v < nV) {
        const uint64_t vis = visit[v];
        const uint64_t nxt = next[v];
        nw = nxt & ~vis & all;
        if (nxt) next[v] = 0;
        if (nw || frontier[v]) frontier[v] = nw;
        if (nw) {
          visit[v] = vis | nw;
          deg = off_out[v + 1] - off_out[v];
          uint64_t bits = nw;
          while (bits) {
            const int s = __builtin_ctzll(bits);
            bits &= bits - 1;
            depth[(size_t)s * nV + v] = level + 1;
          }
        }
      }
      const uint64_t nb = __ballot(nw != 0);
      // frontier bitmap: lanes 0 and 32 write the two words of these vertices
      if ((lane & 31) == 0 && v < nV) fbits[v >> 5] = (uint32_t)(nb >> lane);
      unsigned long long fe = (unsigned long long)deg;
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) fe += __shfl_xor(fe, off, kWave);
      if (lane == 0 && nb) {
        atomicAdd(&ctr->new_vertices[slot_next], (unsigned long long)__popcll(nb));
        atomicAdd(&ctr->frontier_edges[slot_next], fe);
      }
    }
    grid_barrier(grid);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->levels = level + 1;
  }
}

// ------------------------------------------------------------ batched SSSP

__global__ void k_mssp_init(int nV, int ns, const int32_t* __restrict__ sources,
                            int32_t* __restrict__ dist, uint64_t* __restrict__ chg,
                            int32_t inf) {
  const size_t n = (size_t)nV * kWave;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    dist[i] = inf;
    if ((i & 63) == 0) chg[i >> 6] = 0;
  }
}

__global__ void k_mssp_sources(int ns, const int32_t* __restrict__ sources,
                               int32_t* __restrict__ dist, uint64_t* __restrict__ chg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int s = 0; s < ns; ++s) {
    const int v = sources[s];
    dist[(size_t)v * kWave + s] = 0;
    chg[v] |= 1ull << s;
  }
}

// dist (vertex-major [V][64], layout ids) -> out[s * nV + u] in the caller's
// ids (u's layout id is inv[u]): a workgroup moves 64 consecutive caller ids
// through an LDS tile (layout ids to LDS first, then each wave's 16 rows of
// 256 B read with all loads in flight) and writes 64 consecutive values per
// source.
__global__ __launch_bounds__(256) void k_mssp_transpose(int nV, int ns,
                                                        const int32_t* __restrict__ dist,
                                                        const int32_t* __restrict__ inv,
                                                        int32_t* __restrict__ out) {
  constexpr int kRows = kWave / 4;  // rows per wave
  __shared__ int32_t tile[kWave][kWave + 1];
  __shared__ int32_t s_v[kWave];
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  for (int u0 = blockIdx.x * kWave; u0 < nV; u0 += gridDim.x * kWave) {
    if (threadIdx.x < kWave) {
      const int u = u0 + threadIdx.x;
      s_v[threadIdx.x] = u < nV ? (inv ? inv[u] : u) : -1;
    }
    __syncthreads();
    int32_t d[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
      const int v = s_v[w + 4 * k];
      d[k] = v >= 0 ? dist[(size_t)v * kWave + lane] : 0;
    }
#pragma unroll
    for (int k = 0; k < kRows; ++k) tile[w + 4 * k][lane] = d[k];
    __syncthreads();
    for (int s = w; s < ns; s += 4) {
      const int u = u0 + lane;
      if (u < nV) out[(size_t)s * nV + u] = tile[lane][s];
    }
    __syncthreads();
  }
}

// rows[s * nV + v] in layout ids -> out[s * nV + u] in the caller's ids
__global__ void k_rows_to_user(int nV, int ns, const int32_t* __restrict__ rows,
                               const int32_t* __restrict__ inv, int32_t* __restrict__ out) {
  const size_t n = (size_t)nV * ns;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t s = i / nV, u = i % nV;
    out[i] = rows[s * nV + inv[u]];
  }
}

// sources in the caller's ids -> layout ids
__global__ void k_map_ids(int n, const int32_t* __restrict__ in, const int32_t* __restrict__ inv,
                          int32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = inv[in[i]];
}

__global__ void k_permute_i32(const int32_t* __restrict__ in, const int32_t* __restrict__ eid,
                              int64_t n, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[eid ? eid[i] : i];
}

template <class T>
struct DBuf {
  T* p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    if (n < 1) n = 1;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * n);
    if (e != hipSuccess) {
      p = nullptr;
      return set_errorf(GR_ERR_NOMEM, "hipMalloc(%zu bytes): %s", sizeof(T) * n,
                        hipGetErrorString(e));
    }
    return GR_OK;
  }
};

}  // namespace gr

using namespace gr;

// The streaming-pull kernel (k_msbfs_stream) stays off: its cooperative
// launch, at the occupancy query's 8 workgroups per CU, did not finish on an
// MI355X (RMAT-14, 64 sources, automatic direction) although a lane-level
// CPU emulation of the same code terminates with exact depths on that graph.
// k_msbfs did not finish at 8 workgroups per CU either and passes at 4, so
// cooperative grids take at most 4 per CU; whether that was this kernel's
// cause is not established.
constexpr bool kStreamPullEnabled = false;
constexpr int kCoopBlocksPerCu = 4;  // cooperative MS-BFS grid: workgroups per CU (not 8: see above)

// ---------------------------------------------------------- batch context
//
// The state of batched runs on one graph, allocated on the first run that
// needs it and reused by every later run on the context, as an engine binds
// its graph once: the three mask arrays (BFS visit/frontier/next; the first
// two double as SSSP's changed masks), the 1-bit frontier map, the [V][64]
// rows (BFS depths or SSSP distances), the heavy-row chunk tables of the CSC
// and CSR, and for SSSP the weights in CSC order and the recorded rounds.
// Runs on one context are serialised on its stream.
struct gr_batch {
  gr_graph* user_g = nullptr;  // the caller's graph
  gr_graph* g = nullptr;       // its degree-ordered copy (cached in user_g)
  int nV = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool tables = false;
  gr::LaneTables in_chunks, out_chunks;  // heavy CSC rows / CSR rows
  gr::DBuf<uint64_t> m0, m1, m2;
  gr::DBuf<uint32_t> fbits;
  gr::DBuf<int32_t> rows;  // nV * 64
  gr::DBuf<int32_t> d_src, d_src_user;
  gr::DBuf<gr::MsCounters> ctr;
  gr::DBuf<gr::MsLevelCounters> lc;
  void* h_pinned = nullptr;  // host-side counters (pinned)
  // SSSP
  bool has_weights = false;
  gr::DBuf<int32_t> w_csc;
  gr::DBuf<unsigned long long> cnt, any;
  gr::DBuf<uint8_t> dirty;
  gr::DBuf<int32_t> list;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t seg_split = 0;

  ~gr_batch() {
    if (st) (void)hipStreamSynchronize(st);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (h_pinned) (void)hipHostFree(h_pinned);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace gr {
namespace {

constexpr int kSsspCnt = 16;    // counter slots per group (see batch_sssp)
constexpr int kSsspRounds = 8;  // rounds recorded into one hipGraph
static_assert(kSsspRounds <= kSsspCnt, "one counter slot per round");
constexpr size_t kPinnedBytes = 4 * kSsspCnt * sizeof(unsigned long long);
static_assert(kPinnedBytes >= sizeof(MsLevelCounters) && kPinnedBytes >= sizeof(MsCounters),
              "the pinned block holds every counter set");

int batch_common(gr_batch* b) {
  if (b->m0.p) return GR_OK;
  const int nV = b->nV;
  GR_TRY(b->m0.alloc(nV));
  GR_TRY(b->m1.alloc(nV));
  GR_TRY(b->m2.alloc(nV));
  GR_TRY(b->fbits.alloc((size_t)nV / 32 + 2));
  GR_TRY(b->rows.alloc((size_t)nV * kWave));
  GR_TRY(b->d_src.alloc(kWave));
  GR_TRY(b->d_src_user.alloc(kWave));
  GR_TRY(b->ctr.alloc(1));
  GR_TRY(b->lc.alloc(1));
  GR_HIP(hipHostMalloc(&b->h_pinned, kPinnedBytes, hipHostMallocDefault));
  return GR_OK;
}

int batch_tables(gr_batch* b) {
  if (b->tables) return GR_OK;
  GR_TRY(b->in_chunks.build(b->g->h_off_in));
  GR_TRY(b->out_chunks.build(b->g->h_off_out));
  b->tables = true;
  return GR_OK;
}

// The caller's sources, checked, into d_src (layout ids) on the stream.
int batch_sources(gr_batch* b, const char* who, const int32_t* sources, int n_sources) {
  if (!sources || n_sources < 1 || n_sources > 64)
    return set_errorf(GR_ERR_INVALID, "%s: bad arguments", who);
  for (int s = 0; s < n_sources; ++s)
    if (sources[s] < 0 || sources[s] >= b->nV)
      return set_errorf(GR_ERR_INVALID, "%s: source %d out of range", who, sources[s]);
  int32_t* h = static_cast<int32_t*>(b->h_pinned);
  GR_HIP(hipStreamSynchronize(b->st));  // the pinned block is free
  memcpy(h, sources, sizeof(int32_t) * n_sources);
  GR_HIP(hipMemcpyAsync(b->d_src_user.p, h, sizeof(int32_t) * n_sources, hipMemcpyHostToDevice,
                        b->st));
  k_map_ids<<<1, 64, 0, b->st>>>(n_sources, b->d_src_user.p, b->user_g->inv, b->d_src.p);
  GR_HIP(hipGetLastError());
  return GR_OK;
}

int batch_finish(gr_batch* b, int levels, int grid, gr_batch_result* result) {
  GR_HIP(hipEventRecord(b->e1, b->st));
  GR_HIP(hipEventSynchronize(b->e1));
  float ms = 0;
  GR_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
  if (result) {
    result->levels = levels;
    result->ms = ms;
    result->grid = grid;
  }
  return GR_OK;
}

int batch_bfs(gr_batch* b, const int32_t* sources, int n_sources, int32_t* depth_device,
              int direction, double alpha, gr_batch_result* result) {
  const bool stream = (direction & GR_MSBFS_STREAM_PULL) != 0;
  const bool persistent = (direction & GR_MSBFS_PERSISTENT) != 0;
  const bool filter = (direction & GR_MSBFS_FILTER) != 0;
  direction &= ~(GR_MSBFS_STREAM_PULL | GR_MSBFS_PERSISTENT | GR_MSBFS_FILTER);
  if (!depth_device || direction < 0 || direction > 2)
    return set_error(GR_ERR_INVALID, "gr_batch_bfs: bad arguments");
  if (stream && !kStreamPullEnabled)
    return set_error(GR_ERR_INVALID,
                     "gr_batch_bfs: the streaming pull kernel is disabled (its launch did not "
                     "finish on the GPU; cause not established)");
  GR_TRY(batch_common(b));
  GR_TRY(batch_sources(b, "gr_batch_bfs", sources, n_sources));
  GR_TRY(batch_tables(b));
  gr_graph* g = b->g;
  const int nV = b->nV;
  hipStream_t st = b->st;
  uint64_t *visit = b->m0.p, *frontier = b->m1.p, *next = b->m2.p;
  uint32_t* fbits = b->fbits.p;
  int32_t* rows = b->rows.p;
  if (!stream && !persistent) {
    // one expand and one update launch group per level; the host reads the
    // level's two counters through the pinned block
    MsLevelCounters* hlc = static_cast<MsLevelCounters*>(b->h_pinned);
    MsCounters* hc0 = static_cast<MsCounters*>(b->h_pinned);
    GR_HIP(hipEventRecord(b->e0, st));
    // [V][64] depths need no initialisation: the visit bits say which
    // entries are valid
    k_ms_init<<<1024, 256, 0, st>>>(nV, visit, frontier, next, fbits);
    k_ms_sources<<<1, 64, 0, st>>>(n_sources, b->d_src.p, nV, visit, frontier, fbits, rows, 1,
                                   g->off_out, b->ctr.p);
    GR_HIP(hipGetLastError());
    GR_HIP(hipMemcpyAsync(hc0, b->ctr.p, sizeof(MsCounters), hipMemcpyDeviceToHost, st));
    GR_HIP(hipStreamSynchronize(st));
    unsigned long long nv_count = hc0->new_vertices[0], fe_count = hc0->frontier_edges[0];
    const uint64_t all = n_sources >= 64 ? ~0ull : ((1ull << n_sources) - 1);
    const long long threshold = (long long)((double)g->nE / (alpha > 0 ? alpha : 1.0));
    const int grid = (int)std::min<int64_t>(((int64_t)nV + kBatchNT - 1) / kBatchNT, 32768);
    const LaneTables& inc = b->in_chunks;
    const LaneTables& outc = b->out_chunks;
    int levels = 0;
    for (int level = 0; level < nV + 1 && nv_count > 0; ++level) {
      const bool pull = direction == 2 || (direction == 0 && (long long)fe_count > threshold);
      if (pull) {
        k_ms_pull<<<grid, kBatchNT, 0, st>>>(nV, all, g->off_in, g->srcs, visit, frontier,
                                             filter ? fbits : nullptr, next);
        if (inc.n_chunks)
          k_ms_pull_chunks<<<grid_1d(inc.n_chunks, kBatchNT / kWave), kBatchNT, 0, st>>>(
              inc.chunks, inc.n_chunks, all, g->srcs, visit, frontier, next);
      } else {
        k_ms_push<<<grid, kBatchNT, 0, st>>>(nV, g->off_out, g->dsts, visit, frontier, fbits,
                                             next);
        if (outc.n_chunks)
          k_ms_push_chunks<<<grid_1d(outc.n_chunks, kBatchNT / kWave), kBatchNT, 0, st>>>(
              outc.chunks, outc.n_chunks, g->dsts, visit, frontier, fbits, next);
      }
      GR_HIP(hipMemsetAsync(b->lc.p, 0, sizeof(MsLevelCounters), st));
      // (two counter adds per workgroup)
      k_ms_update<<<std::min(grid, kAtomicGrid), kBatchNT, 0, st>>>(nV, all, level, g->off_out, visit, frontier, next,
                                             fbits, rows, b->lc.p);
      GR_HIP(hipGetLastError());
      GR_HIP(hipMemcpyAsync(hlc, b->lc.p, sizeof(MsLevelCounters), hipMemcpyDeviceToHost, st));
      GR_HIP(hipStreamSynchronize(st));
      nv_count = hlc->new_vertices;
      fe_count = hlc->frontier_edges;
      levels = level + 1;
    }
    k_ms_depth_out<<<(unsigned)std::min<int64_t>(((int64_t)nV + kWave - 1) / kWave, 65536), 256,
                     0, st>>>(nV, n_sources, rows, visit, b->user_g->inv, depth_device);
    GR_HIP(hipGetLastError());
    return batch_finish(b, levels, grid, result);
  }
  const void* kernel = stream ? (const void*)k_msbfs_stream : (const void*)k_msbfs;
  int dev = 0, cus = 0, per_cu = 0;
  GR_HIP(hipGetDevice(&dev));
  GR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  GR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBatchNT, 0));
  if (per_cu < 1) return set_error(GR_ERR_HIP, "gr_batch_bfs: kernel cannot be resident");
  // Every workgroup must be resident at once (grid.sync waits for all of
  // them).  The compiler's resource report for gfx950 (k_msbfs: 100 SGPRs,
  // 44 VGPRs, 64 B of LDS; k_msbfs_stream: 92 SGPRs, 48 VGPRs, no LDS) gives
  // 8 waves per SIMD for both, i.e. 8 workgroups of 4 waves per CU, which is
  // what the occupancy query returns.  The grid takes at most half of that,
  // so the launch stays resident even when other work holds part of every
  // CU; the level loops are grid-strided and do not depend on the grid size.
  const int grid = cus * std::max(1, std::min(per_cu, kCoopBlocksPerCu));
  GR_HIP(hipEventRecord(b->e0, st));
  int nVa = nV, ns = n_sources, mode = direction, max_levels = nV + 1;
  long long nE = g->nE;
  double al = alpha;
  const int32_t *off_out = g->off_out, *dsts = g->dsts, *off_in = g->off_in, *srcs = g->srcs;
  uint64_t *pv = visit, *pf = frontier, *pn = next;
  uint32_t* pfb = fbits;
  int32_t* pd = rows;
  MsCounters* pc = b->ctr.p;
  k_ms_init<<<1024, 256, 0, st>>>(nV, visit, frontier, next, fbits);
  if (stream) {
    // the streaming kernel writes [ns][V] rows: depth = -1 everywhere first,
    // the counterpart of the host-side vertex init before setGraph
    GR_HIP(hipMemsetAsync(rows, 0xFF, sizeof(int32_t) * (size_t)nV * n_sources, st));
    k_ms_sources<<<1, 64, 0, st>>>(n_sources, b->d_src.p, nV, visit, frontier, fbits, rows, 0,
                                   g->off_out, b->ctr.p);
    GR_HIP(hipGetLastError());
    const LaneChunk* pch = b->in_chunks.chunks;
    int nch = b->in_chunks.n_chunks;
    void* args[] = {&nVa, &nE, &ns, &mode, &al, &off_out, &dsts, &off_in, &srcs, &pv,
                    &pf,  &pn, &pfb, &pd,  &pch, &nch,   &pc,   &max_levels};
    GR_HIP(hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(kBatchNT), args, 0, st));
    k_rows_to_user<<<4096, 256, 0, st>>>(nV, n_sources, rows, b->user_g->inv, depth_device);
  } else {
    // vertex-major [V][64] depths, valid where the visit bit is set
    k_ms_sources<<<1, 64, 0, st>>>(n_sources, b->d_src.p, nV, visit, frontier, fbits, rows, 1,
                                   g->off_out, b->ctr.p);
    GR_HIP(hipGetLastError());
    int flt = filter ? 1 : 0;
    const LaneChunk *pin = b->in_chunks.chunks, *pout = b->out_chunks.chunks;
    int nin = b->in_chunks.n_chunks, nout = b->out_chunks.n_chunks;
    void* args[] = {&nVa, &nE,  &ns,  &mode, &al,  &flt, &off_out, &dsts, &off_in, &srcs, &pv,
                    &pf,  &pn,  &pfb, &pd,   &pin, &nin, &pout,    &nout, &pc,     &max_levels};
    GR_HIP(hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(kBatchNT), args, 0, st));
    k_ms_depth_out<<<(unsigned)std::min<int64_t>(((int64_t)nV + kWave - 1) / kWave, 65536), 256,
                     0, st>>>(nV, n_sources, rows, visit, b->user_g->inv, depth_device);
  }
  GR_HIP(hipGetLastError());
  MsCounters* hc = static_cast<MsCounters*>(b->h_pinned);
  GR_HIP(hipMemcpyAsync(hc, b->ctr.p, sizeof(MsCounters), hipMemcpyDeviceToHost, st));
  GR_TRY(batch_finish(b, 0, grid, result));
  if (result) result->levels = hc->levels;
  return GR_OK;
}

// The SSSP state and the recorded rounds, on the first SSSP run of a context.
int batch_sssp_setup(gr_batch* b) {
  if (b->exec) return GR_OK;
  gr_graph* g = b->g;
  const int nV = b->nV;
  hipStream_t st = b->st;
  // per round r of a replay: cnt[r] changed vertices, cnt[kSsspCnt + r] their
  // out-edges, cnt[2 kSsspCnt + r] filtered-list length; cnt[3 kSsspCnt +
  // 0/1]: the previous replay's last round (changed vertices, out-edges)
  GR_TRY(b->cnt.alloc(4 * kSsspCnt));
  GR_TRY(b->any.alloc(1));
  GR_TRY(b->dirty.alloc(nV));
  GR_TRY(b->list.alloc(nV));
  b->seg_split = graph_lane_split(g);
  const LaneTables& lt = b->in_chunks;
  const LaneTables& lt_out = b->out_chunks;
  int32_t* dist = b->rows.p;
  // the filter kernel adds its count once per workgroup; the light kernel
  // too, but it runs faster on a wave per vertex than with fewer workgroups
  // (measured at RMAT-24: 65,536 workgroups 68 ms per batch, 2,048 83 ms)
  const unsigned rgrid = std::min<unsigned>(grid_1d((int64_t)nV, kLaneNT / kWave), 65536);
  const unsigned hgrid = grid_1d((int64_t)lt.n_chunks, kLaneNT / kWave);
  const unsigned vgrid = std::min<unsigned>(grid_1d((int64_t)nV, kLaneNT), kAtomicGrid);
  // a filtered round pays one flag write per out-edge of the changed vertices;
  // it replaces the scan of every in-edge when those are a small share
  const long long filter_max = g->nE / 8;
  unsigned long long* cnt = b->cnt.p;
  GR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int r = 0; r < kSsspRounds; ++r) {
    uint64_t* cin = (r & 1) ? b->m1.p : b->m0.p;
    uint64_t* cout = (r & 1) ? b->m0.p : b->m1.p;
    // a round after a round without change returns at once (exact: nothing
    // changed feeds it), so the rest of a replay after convergence is free
    LaneRound rr;
    rr.prev_count = r ? cnt + (r - 1) : cnt + 3 * kSsspCnt;
    rr.prev_edges = r ? cnt + kSsspCnt + (r - 1) : cnt + 3 * kSsspCnt + 1;
    rr.count = cnt + r;
    rr.edges = cnt + kSsspCnt + r;
    rr.dirty = b->dirty.p;
    rr.list = b->list.p;
    rr.list_len = reinterpret_cast<unsigned int*>(cnt + 2 * kSsspCnt + r);
    rr.filter_max = filter_max;
    rr.host_filter = -1;
    k_lane_sssp_mark<<<vgrid, kLaneNT, 0, st>>>(nV, g->off_out, g->dsts, lt_out.chunks,
                                                lt_out.n_chunks, cin, rr);
    k_lane_sssp_filter<<<vgrid, kLaneNT, 0, st>>>(nV, g->off_in, cout, rr);
    k_lane_sssp_light<<<rgrid, kLaneNT, 0, st>>>(
        nV, g->off_in, g->off_out, g->srcs, b->w_csc.p, dist, cin, cout, b->any.p, rr,
        b->seg_split);
    if (lt.n_chunks) {
      k_lane_sssp_heavy<<<hgrid, kLaneNT, 0, st>>>(lt.chunks, lt.n_chunks, g->off_out, g->srcs,
                                                   b->w_csc.p, dist, cin, cout, b->any.p, rr);
      k_lane_sssp_clear_heavy<<<grid_1d(lt.n_heavy, 256), 256, 0, st>>>(lt.heavy, lt.n_heavy,
                                                                        rr);
    }
  }
  GR_HIP(hipStreamEndCapture(st, &b->graph));
  GR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return GR_OK;
}

int batch_sssp(gr_batch* b, const int32_t* sources, int n_sources, int32_t* dist_device,
               gr_batch_result* result) {
  if (!dist_device) return set_error(GR_ERR_INVALID, "gr_batch_sssp: bad arguments");
  if (!b->has_weights)
    return set_error(GR_ERR_INVALID, "gr_batch_sssp: the context was created without weights");
  const int32_t inf = 2147483647 - 100000;  // SSSP::gatherZero, sssp.cu:32-33
  GR_TRY(batch_common(b));
  GR_TRY(batch_sources(b, "gr_batch_sssp", sources, n_sources));
  GR_TRY(batch_tables(b));
  GR_TRY(batch_sssp_setup(b));
  const int nV = b->nV;
  hipStream_t st = b->st;
  int32_t* dist = b->rows.p;
  GR_HIP(hipEventRecord(b->e0, st));
  // the flags are all zero between rounds of a run; cleared here so that a
  // run never depends on how the previous one ended
  GR_HIP(hipMemsetAsync(b->dirty.p, 0, (size_t)nV, st));
  GR_HIP(hipMemsetAsync(b->any.p, 0, sizeof(unsigned long long), st));
  k_mssp_init<<<4096, 256, 0, st>>>(nV, n_sources, b->d_src.p, dist, b->m0.p, inf);
  k_mssp_sources<<<1, 64, 0, st>>>(n_sources, b->d_src.p, dist, b->m0.p);
  GR_HIP(hipGetLastError());
  unsigned long long* hall = static_cast<unsigned long long*>(b->h_pinned);
  int rounds = 0;
  // the first replay's first round has no previous round: not idle, unfiltered
  unsigned long long carry[2] = {1ull, ~0ull};
  for (;;) {
    // counters are cleared by a kernel-free copy from the pinned block, and
    // the last round's counts carried into the next replay's first round
    GR_HIP(hipStreamSynchronize(st));  // the pinned block is free
    memset(hall, 0, kPinnedBytes);
    hall[3 * kSsspCnt] = carry[0];
    hall[3 * kSsspCnt + 1] = carry[1];
    GR_HIP(hipMemcpyAsync(b->cnt.p, hall, kPinnedBytes, hipMemcpyHostToDevice, st));
    GR_HIP(hipGraphLaunch(b->exec, st));
    GR_HIP(hipMemcpyAsync(hall, b->cnt.p, kPinnedBytes, hipMemcpyDeviceToHost, st));
    GR_HIP(hipStreamSynchronize(st));
    carry[0] = hall[kSsspRounds - 1];
    carry[1] = hall[kSsspCnt + kSsspRounds - 1];
    int last = -1;
    for (int r = 0; r < kSsspRounds; ++r)
      if (hall[r]) last = r;
    if (last < kSsspRounds - 1) {
      rounds += last + 1;
      break;
    }
    rounds += kSsspRounds;
  }
  k_mssp_transpose<<<(unsigned)std::min<int64_t>(((int64_t)nV + kWave - 1) / kWave, 65536), 256, 0,
                     st>>>(nV, n_sources, dist, b->user_g->inv, dist_device);
  GR_HIP(hipGetLastError());
  return batch_finish(b, rounds, (int)grid_1d((int64_t)nV, kLaneNT / kWave), result);
}

}  // namespace
}  // namespace gr

extern "C" int gr_batch_create(gr_graph* g, const int32_t* weights_device, gr_batch** out) {
  if (!g || !out) return set_error(GR_ERR_INVALID, "gr_batch_create: bad arguments");
  *out = nullptr;
  gr_batch* b = new (std::nothrow) gr_batch;
  if (!b) return set_error(GR_ERR_NOMEM, "gr_batch_create: out of host memory");
  b->user_g = g;
  b->nV = g->nV;
  int rc = graph_ordered(g, &b->g);
  if (rc == GR_OK && hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
    rc = set_error(GR_ERR_HIP, "gr_batch_create: hipStreamCreateWithFlags failed");
  if (rc == GR_OK &&
      (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess))
    rc = set_error(GR_ERR_HIP, "gr_batch_create: hipEventCreate failed");
  if (rc == GR_OK && weights_device) {
    // the weights in the CSC order of the degree-ordered copy
    b->has_weights = true;
    rc = b->w_csc.alloc((size_t)b->g->nE);
    if (rc == GR_OK && b->g->nE) {
      k_permute_i32<<<grid_1d(b->g->nE, 256), 256, 0, b->st>>>(weights_device, b->g->csc_eid,
                                                              b->g->nE, b->w_csc.p);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->st) != hipSuccess)
        rc = set_error(GR_ERR_HIP, "gr_batch_create: weight permutation failed");
    }
  }
  if (rc != GR_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return GR_OK;
}

extern "C" int gr_batch_bfs(gr_batch* b, const int32_t* sources, int n_sources,
                            int32_t* depth_device, int direction, double alpha,
                            gr_batch_result* result) {
  if (!b) return set_error(GR_ERR_INVALID, "gr_batch_bfs: bad arguments");
  return batch_bfs(b, sources, n_sources, depth_device, direction, alpha, result);
}

extern "C" int gr_batch_sssp(gr_batch* b, const int32_t* sources, int n_sources,
                             int32_t* dist_device, gr_batch_result* result) {
  if (!b) return set_error(GR_ERR_INVALID, "gr_batch_sssp: bad arguments");
  return batch_sssp(b, sources, n_sources, dist_device, result);
}

extern "C" int gr_batch_destroy(gr_batch* b) {
  delete b;
  return GR_OK;
}

// One-shot forms: a context for one run.
extern "C" int gr_msbfs(gr_graph* g, const int32_t* sources, int n_sources, int32_t* depth_device,
                        int direction, double alpha, gr_batch_result* result) {
  if (!g || !sources || !depth_device || n_sources < 1 || n_sources > 64)
    return set_error(GR_ERR_INVALID, "gr_msbfs: bad arguments");
  gr_batch* b = nullptr;
  GR_TRY(gr_batch_create(g, nullptr, &b));
  const int rc = batch_bfs(b, sources, n_sources, depth_device, direction, alpha, result);
  delete b;
  return rc;
}

extern "C" int gr_mssp(gr_graph* g, const int32_t* weights_device, const int32_t* sources,
                       int n_sources, int32_t* dist_device, gr_batch_result* result) {
  if (!g || !weights_device || !sources || !dist_device || n_sources < 1 || n_sources > 64)
    return set_error(GR_ERR_INVALID, "gr_mssp: bad arguments");
  gr_batch* b = nullptr;
  GR_TRY(gr_batch_create(g, weights_device, &b));
  const int rc = batch_sssp(b, sources, n_sources, dist_device, result);
  delete b;
  return rc;
}
