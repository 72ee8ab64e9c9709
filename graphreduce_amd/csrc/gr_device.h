// Device-side building blocks shared by the HIP sources: error checks, wave64
// primitives, and the device graph object.
#ifndef GRAPHREDUCE_AMD_GR_DEVICE_H
#define GRAPHREDUCE_AMD_GR_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "gr_internal.h"

#define GR_HIP(expr)                                                               \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return ::gr::set_errorf(GR_ERR_HIP, "%s:%d: %s failed: %s", __FILE__, __LINE__, \
                              #expr, hipGetErrorString(_e));                       \
  } while (0)

#define GR_TRY(expr)            \
  do {                          \
    int _rc = (expr);           \
    if (_rc != GR_OK) return _rc; \
  } while (0)

namespace gr {

constexpr int kWave = 64;

// ------------------------------------------------------------ wave64 helpers

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Value of `x` in lane `l` as a wave-uniform (scalar) value; v_readlane
// ignores the exec mask, so the lane need not be active.  Loop bounds read
// this way keep a loop scalar: a bound the compiler must treat as per-lane
// lets it run lanes through different iterations, in which a shuffle from a
// lane that is not running reads 0.
__device__ __forceinline__ int lane_value(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ uint64_t lane_value64(uint64_t x, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

template <class T>
__device__ __forceinline__ T wave_incl_scan_add(T v) {
  const int l = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T o = __shfl_up(v, off, kWave);
    if (l >= off) v += o;
  }
  return v;
}

// Inclusive prefix sum of an int over the wave by DPP moves, without the LDS
// crossbar that shuffles use: shifts by 1, 2, 4 and 8 lanes within each row of
// 16 lanes (lanes whose source lies before their row add 0), then the last
// lane of row 0 into row 1 and of row 2 into row 3 (row_bcast:15, rows 1 and
// 3 written), then lane 31 into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ int wave_incl_scan_add_dpp(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

template <class T>
__device__ __forceinline__ T wave_reduce_add(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Exclusive prefix sum over a workgroup of NT threads (NT a multiple of 64).
// `scratch` must hold NT / 64 elements.  Returns the exclusive prefix and the
// block total in *total.
template <int NT, class T>
__device__ __forceinline__ T block_excl_scan_add(T v, T* scratch, T* total) {
  constexpr int NW = NT / kWave;
  const int l = lane_id(), w = threadIdx.x / kWave;
  T incl = wave_incl_scan_add(v);
  if (l == kWave - 1) scratch[w] = incl;
  __syncthreads();
  T before = 0, sum = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    T s = scratch[i];
    if (i < w) before += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + incl - v;
}

// ----------------------------------------------------------- device graph

}  // namespace gr

// Row blocks of the CSC whose in-edges are listed by source (for the blocked
// gather): block b holds rows [brow[b], brow[b] + bnr[b]) -- at most
// gr::kBlockRows (or gr::kBlockRowsWide) rows and, unless one row alone has more, the edge target
// below -- and its in-edges, sorted by source (stable), fill slots
// [boff[b], boff[b + 1]) of the packed layout.  Rows from rows_end on have no
// in-edges and belong to no block.  The edge target per block is
// E / (2 * compute units), at most kBlockEdges and at least 4096.
//
// Packed layout: slot s holds bpk[s] = step << 16 | row, where row is the
// in-edge's row - brow[b] and step its source minus the source of slot s - 1.
// Slots come in groups of gr::kPackGroup; the first slot of a group stores
// step 0 and gbase[group] holds its source, so that a source is gbase plus a
// prefix sum of steps within the group.  Where two consecutive in-edges of a
// block have sources more than gr::kPackStep apart, pad slots (row
// gr::kPackPad) of step kPackStep come between them; every block starts on a
// group and its last group is filled up with pad slots of step 0.  bpos[s] is
// the CSC position of slot s's in-edge (-1 for a pad slot), which lays the
// edge data out in slot order.
struct gr_blocked {
  int32_t n_blocks = 0;
  int32_t rows_end = 0;
  int64_t n_slots = 0;      // a multiple of kPackGroup
  int32_t* brow = nullptr;  // [n_blocks]
  int32_t* bnr = nullptr;   // [n_blocks]
  int32_t* boff = nullptr;  // [n_blocks + 1] first slot of each block
  uint32_t* bpk = nullptr;  // [n_slots] step << 16 | row
  int32_t* gbase = nullptr; // [n_slots / kPackGroup] source of each group's first slot
  int32_t* bpos = nullptr;  // [n_slots] CSC position, -1 for a pad slot
};

struct gr_graph {
  int32_t nV = 0;
  int64_t nE = 0;
  int32_t* off_out = nullptr;  // CSR offsets [V+1]
  int32_t* dsts = nullptr;     // CSR targets [E]
  int32_t* off_in = nullptr;   // CSC offsets [V+1]
  int32_t* srcs = nullptr;     // CSC sources [E]
  int32_t* csr_eid = nullptr;  // CSR position -> edge-list index (nullptr: identity)
  int32_t* csc_eid = nullptr;  // CSC position -> edge-list index
  std::vector<int32_t> h_off_out, h_off_in;  // host copies of the offsets
  // Degree-ordered copy for the engines, built on first use: vertex v of
  // `ordered` is vertex perm[v] of this graph (descending in+out degree,
  // ties by id); its rows are sorted by the new neighbour ids and its eid
  // maps index this graph's edge list.
  gr_graph* ordered = nullptr;
  // 1 when the degrees are skewed enough for the degree-ordered copy to pay
  // (graph_skewed; -1 until first asked)
  int skewed = -1;
  // first vertex of the lane kernels' segment walk (lane_seg_split; -1 until
  // first asked)
  int32_t lane_split = -1;
  int32_t* perm = nullptr;  // ordered id -> id   [V], device
  int32_t* inv = nullptr;   // id -> ordered id   [V], device
  // row blocks of at most kBlockRows / kBlockRowsWide rows, built on first
  // use (graph_blocked)
  gr_blocked* blocked = nullptr;
  gr_blocked* blocked_half = nullptr;
};

namespace gr {
// Grid-strided kernels whose workgroups each add to one counter take at most
// this many workgroups (8 per CU of 4 waves: every CU full): device-scope
// atomics on one address are performed one after another at the memory side,
// about 12 ns each, so a launch of 65,536 such workgroups would spend
// 0.8 ms on them.
constexpr int kAtomicGrid = 2048;
constexpr int kBlockRows = 32768;       // rows per gather block (LDS: 128 KB of int)
constexpr int kBlockRowsWide = 16384;   // ... for 8-byte accumulators (LDS: 128 KB)
#ifndef GR_MIN_BLOCK_ROWS
#define GR_MIN_BLOCK_ROWS kBlockRows
#endif
constexpr int kMinBlockRows = GR_MIN_BLOCK_ROWS;  // rows per block of the int-minimum gather
// ... on a graph of uniform degree (graph_skewed false): two workgroups per
// CU, since there the source reads mostly hit the cache and a workgroup's
// passes over its slots, one after another, are what take the time
#ifndef GR_MIN_BLOCK_ROWS_FLAT
#define GR_MIN_BLOCK_ROWS_FLAT kBlockRowsWide
#endif
constexpr int kMinBlockRowsFlat = GR_MIN_BLOCK_ROWS_FLAT;
static_assert((kMinBlockRows == kBlockRows || kMinBlockRows == kBlockRowsWide) &&
                  (kMinBlockRowsFlat == kBlockRows || kMinBlockRowsFlat == kBlockRowsWide),
              "the row blocks are built for these two limits");
constexpr int kBlockEdges = 1 << 19;    // in-edges per gather block (largest target)
constexpr int kPackGroup = 512;          // slots per group of the packed layout
constexpr uint32_t kPackPad = 0xFFFFu;   // row of a pad slot
constexpr int32_t kPackStep = 0xFFFF;    // largest source step one slot holds
static_assert(kBlockRows <= (int)kPackPad && kBlockRowsWide <= (int)kPackPad,
              "a block's rows fit below the pad row");
int graph_free(gr_graph* g);
// The degree-ordered copy of g (cached in g; freed with g).
int graph_ordered(gr_graph* g, gr_graph** out);
// Whether some vertex's in+out degree is at least kSkewFactor times the mean
// (2E / V): then the hubs' values are worth packing into few cache lines by
// the degree order.  A graph of uniform degree (a grid, a road map) keeps its
// own numbering, whose neighbourhoods are already local.
constexpr int kSkewFactor = 32;
inline bool graph_skewed(gr_graph* g) {
  if (g->skewed < 0 && (g->h_off_out.size() != (size_t)g->nV + 1 ||
                         g->h_off_in.size() != (size_t)g->nV + 1))
    return true;  // no host offsets to look at: keep the degree order
  if (g->skewed < 0) {
    int64_t mx = 0;
    for (int32_t v = 0; v < g->nV; ++v) {
      const int64_t d = (int64_t)(g->h_off_out[v + 1] - g->h_off_out[v]) +
                        (int64_t)(g->h_off_in[v + 1] - g->h_off_in[v]);
      if (d > mx) mx = d;
    }
    g->skewed = g->nV > 0 && mx * (int64_t)g->nV >= (int64_t)kSkewFactor * 2 * g->nE ? 1 : 0;
  }
  return g->skewed == 1;
}
// The row blocks of g's CSC with in-edges listed by source, at most max_rows
// (kBlockRows or kBlockRowsWide) rows each (cached in g).
int graph_blocked(gr_graph* g, int max_rows, const gr_blocked** out);
}

#endif
