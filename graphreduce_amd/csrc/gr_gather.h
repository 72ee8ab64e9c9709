// Merge-path gather: the load-balanced, segmented reduction of a vertex
// program's gatherMap values over the CSC in-edges of an active list.  Shared
// by the GAS engine (gr_engine.hip) and the PageRank precision form
// (gr_pagerank.hip).
//
// Work is the merge of the active rows' end offsets (exclusive scan `escan`
// of their in-degrees) with the edge items; fixed tiles of kNT * VT path items
// go to one workgroup each (k_merge_partition finds the tile starts).  Inside
// a tile: (1) each lane walks its VT items and labels edges with rows,
// (2) a coalesced pass reads the CSC sources and edge data and applies
// gatherMap into LDS, (3) each lane reduces its items sequentially and a
// wave64 segmented scan joins the lanes' partial rows.  Rows that span tiles
// leave a carry and a head partial, joined in tile order by k_gather_fixup,
// so results are deterministic.
#ifndef GRAPHREDUCE_AMD_GR_GATHER_H
#define GRAPHREDUCE_AMD_GR_GATHER_H

#include <type_traits>

#include "gr_device.h"

namespace gr {

constexpr int kNT = 256;            // threads per tile workgroup (4 waves)
constexpr int kVT = 8;              // merge-path items per thread (push tiles)
constexpr int kTile = kNT * kVT;    // path items per push tile
constexpr int kCompactNT = 256;     // compaction: one bitmap word per thread
constexpr int kMaxGrid = 1 << 20;

struct DevCounters {
  int32_t n_active;               // size of the active list built by nextIter
  int32_t w_edges;                // sum of in-degrees over that list
  unsigned long long chg;         // (#changed << 32) | out-edges, appended by apply
  unsigned long long chg_last;    // chg as seen by the last nextIter
  int32_t pull;                   // last scatterActivate ran in pull mode
  int32_t fused;                  // last scatterActivate also pushed the next gather
  int32_t fab;                    // last scatterActivate ran the next blocked gather
  unsigned long long inert;       // (#non-inert << 32) | out-edges, from k_inert_scan
  unsigned long long applied_last;  // vertices whose apply returned true since
                                    // the nextIter before the last one
};

// Vertices whose apply returned true (with or without out-edges) are counted
// by apply per workgroup into kAppliedSlots counters 128 bytes apart (slot =
// workgroup id mod kAppliedSlots), so that the adds do not contend on one
// address; nextIter sums and clears them.
constexpr int kAppliedSlots = 64;
constexpr int kAppliedStride = 16;  // unsigned long longs between slots
static_assert(kAppliedSlots <= kWave, "nextIter sums the slots in one wave");

// Read-once operands (CSC sources, edge data) are loaded non-temporally so
// that streaming them does not displace the randomly re-read vertex values
// from L2.
template <class T>
__device__ __forceinline__ T stream_load(const T* p) {
  if constexpr (std::is_empty<T>::value) {
    return T();
  } else if constexpr (sizeof(T) % 4 == 0) {
    T out;
    uint32_t* o = reinterpret_cast<uint32_t*>(&out);
    const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = __builtin_nontemporal_load(q + i);
    return out;
  } else {
    return *p;
  }
}

template <class T>
struct Part {
  int32_t row;
  int32_t has;
  T val;
};

static inline unsigned grid_1d(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return (unsigned)g;
}


// Shuffle of a plain-old-data value, 32 bits at a time.
template <class T>
__device__ __forceinline__ T shfl_up_pod(const T& v, int off) {
  static_assert(sizeof(T) % 4 == 0, "shuffled values must be a multiple of 4 bytes");
  T r;
  const int* a = reinterpret_cast<const int*>(&v);
  int* b = reinterpret_cast<int*>(&r);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) b[i] = __shfl_up(a[i], off, kWave);
  return r;
}

// ------------------------------------------------------------ merge path

// Number of row ends consumed on diagonal d of the merge of rowEnd[0..n)
// (inclusive-scan row ends) with the edge items [0, w).  Edge j belongs to row
// i while j < rowEnd[i].
__device__ __forceinline__ int merge_path_search(int d, const int32_t* __restrict__ rowEnd,
                                                 int n, int w) {
  int lo = max(d - w, 0), hi = min(d, n);
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (rowEnd[mid] <= d - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// (static: every translation unit that launches it has its own copy)
[[maybe_unused]] static __global__ void k_merge_partition(const int32_t* __restrict__ escan, int n,
                                                          int w, int n_tiles, int tile,
                                                          int32_t* __restrict__ part) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  int64_t d = min((int64_t)t * tile, (int64_t)n + w);
  part[t] = merge_path_search((int)d, escan + 1, n, w);
}

// ------------------------------------------------- segmented lane scan

template <class P>
struct Seg {
  typedef typename P::GatherResult G;
  int flag;  // this lane emitted a row (its tail starts a new row)
  int has;
  G val;

  __device__ static Seg identity() {
    Seg s;
    s.flag = 0;
    s.has = 0;
    s.val = G();
    return s;
  }
  // a precedes b on the merge path
  __device__ static Seg combine(const Seg& a, const Seg& b) {
    Seg r;
    r.flag = a.flag | b.flag;
    if (b.flag || !a.has) {
      r.has = b.has;
      r.val = b.val;
    } else if (!b.has) {
      r.has = a.has;
      r.val = a.val;
    } else {
      r.has = 1;
      r.val = P::gatherReduce(a.val, b.val);
    }
    return r;
  }
  __device__ static Seg shfl_up(const Seg& s, int off) {
    Seg r;
    r.flag = __shfl_up(s.flag, off, kWave);
    r.has = __shfl_up(s.has, off, kWave);
    r.val = shfl_up_pod(s.val, off);
    return r;
  }
};

// Delta gather: a program opts in with kDeltaGather = true when gatherReduce
// is min, gatherMap depends only on the source value and the edge, and apply
// keeps min(current, g) and reports exactly a decrease.  Programs without the
// member are treated as false.
template <class P, class = void>
struct delta_gather : std::false_type {};
template <class P>
struct delta_gather<P, std::void_t<decltype(P::kDeltaGather)>>
    : std::integral_constant<bool, P::kDeltaGather> {};
template <class P>
constexpr bool delta_gather_v = delta_gather<P>::value;

// Programs whose gatherMap reads the source only through a smaller value
// declare SourceValue, sourceValue(vertex) and gatherMapValue(value, edge):
// the engine keeps that value per vertex and the gather reads it instead of
// the whole vertex data (PageRank: rank / numOutEdges, 4 B instead of 8 B per
// random source read, and one division per vertex instead of per edge).
template <class P, class = void>
struct source_value {
  static constexpr bool value = false;
  typedef int type;
};
template <class P>
struct source_value<P, std::void_t<typename P::SourceValue>> {
  static constexpr bool value = true;
  typedef typename P::SourceValue type;
};
template <class P>
constexpr bool source_value_v = source_value<P>::value;
template <class P>
using source_value_t = typename source_value<P>::type;

// Position of the edge data of CSC slot c: programs with a scatter keep their
// edge data in edge-list order (one copy that the CSR-ordered scatter writes
// and the gather reads through csc_eid, as refgas.h:150-152 and 192-196 do),
// the others keep a CSC-ordered copy for coalesced reads.
template <class P>
__device__ __forceinline__ int edge_pos(const int32_t* __restrict__ ed_index, int c) {
  if constexpr (P::kHasScatter)
    return ed_index[c];
  else
    return c;
}

// Programs whose gather is a sum of non-negative floats may declare
// kFixedPointSum: the blocked gather then adds their terms exactly in 64-bit
// fixed point (each term rounded once to a multiple of 2^-32) and rounds the
// row's sum once to float, so the result does not depend on the order in
// which the LDS additions land (determinism) and is at least as close to the
// exact sum as a float tree sum.
template <class P, class = void>
struct fixed_point_sum : std::false_type {};
template <class P>
struct fixed_point_sum<P, std::void_t<decltype(P::kFixedPointSum)>>
    : std::integral_constant<bool, P::kFixedPointSum> {};
template <class P>
constexpr bool fixed_point_sum_v = fixed_point_sum<P>::value;

// Programs with kFixedPointSum whose kept source value is a float must keep
// that value non-negative (PageRank's rank / out-degree is), so its sign bit
// is free: the
// engine may keep there whether the vertex's last apply returned true and it
// has out-edges (flagged(), unflag()), which the fused blocked sum then reads
// with the value instead of the changed bitmap.  Every other reader masks it.
template <class P>
constexpr bool source_flag_v =
    fixed_point_sum_v<P> && source_value_v<P> && std::is_same<source_value_t<P>, float>::value;
__device__ __forceinline__ float unflag(float s) {
  return __uint_as_float(__float_as_uint(s) & 0x7FFFFFFFu);
}
__device__ __forceinline__ float flagged(float s) {
  return __uint_as_float(__float_as_uint(s) | 0x80000000u);
}

// gatherMap of a source, through its kept source value when there is one.
template <class P>
__device__ __forceinline__ typename P::GatherResult map_source(
    const typename P::VertexData* dst, const typename P::VertexData* vd,
    const source_value_t<P>* srcval, int src, const typename P::EdgeData* edge) {
  if constexpr (source_value_v<P>) {
    if (srcval) {
      if constexpr (source_flag_v<P>)
        return P::gatherMapValue(unflag(srcval[src]), edge);
      else
        return P::gatherMapValue(srcval[src], edge);
    }
  }
  return P::gatherMap(dst, vd + src, edge);
}

// Programs with kInertSources name vertex values whose gatherMap contribution
// gatherZero absorbs on every edge whose data passes P::edgeKeepsInert.
template <class P, class = void>
struct inert_sources : std::false_type {};
template <class P>
struct inert_sources<P, std::void_t<decltype(P::kInertSources)>>
    : std::integral_constant<bool, P::kInertSources> {};
template <class P>
constexpr bool inert_sources_v = inert_sources<P>::value;

template <class P>
__device__ __forceinline__ typename P::GatherResult finalize(int has,
                                                             typename P::GatherResult v) {
  return has ? P::gatherReduce(P::gatherZero, v) : P::gatherZero;
}

// ------------------------------------------------------------------ gather

template <class P, int VT>
__global__ __launch_bounds__(kNT) void k_gather_tiles(
    int n, int w, const int32_t* __restrict__ active, const int32_t* __restrict__ escan,
    const int32_t* __restrict__ part, const int32_t* __restrict__ cbase,
    const int32_t* __restrict__ srcs, const typename P::VertexData* __restrict__ vd,
    const typename P::EdgeData* __restrict__ ed, typename P::GatherResult* __restrict__ gtmp,
    Part<typename P::GatherResult>* __restrict__ carry,
    Part<typename P::GatherResult>* __restrict__ head,
    const uint32_t* __restrict__ delta_bits, const source_value_t<P>* __restrict__ srcval,
    const int32_t* __restrict__ ed_index) {
  typedef typename P::GatherResult G;
  typedef Seg<P> S;
  constexpr int TILE = kNT * VT;
  __shared__ int32_t s_scan[TILE + 2];
  __shared__ int32_t s_base[TILE + 1];
  __shared__ uint16_t s_row[TILE];
  __shared__ G s_val[TILE];
  __shared__ S s_wave[kNT / kWave];

  const int t = blockIdx.x;
  const int tid = threadIdx.x;
  const int total = n + w;
  const int d0 = t * TILE;
  const int d1 = min(d0 + TILE, total);
  const int i0 = part[t], i1 = part[t + 1];
  const int j0 = d0 - i0, j1 = d1 - i1;
  const int n_rows = i1 - i0;                    // row ends inside this tile
  const int n_ref = n_rows + (i1 < n ? 1 : 0);   // plus the row still open at the end
  const int n_edges = j1 - j0;

  for (int k = tid; k <= n_rows; k += kNT) s_scan[k] = escan[i0 + k];
  // cbase == nullptr: every row's CSC base is 0 (the active list is every
  // vertex in order and escan is the CSC offset array itself)
  for (int k = tid; k < n_ref; k += kNT) s_base[k] = cbase ? cbase[i0 + k] : 0;
  __syncthreads();

  // Phase 1: each lane walks its VT path items and labels edges with rows.
  const int dl = tid * VT;
  const int steps = max(0, min(VT, (d1 - d0) - dl));
  int ri = 0, ej = 0;
  if (steps > 0) {
    int lo = max(dl - n_edges, 0), hi = min(dl, n_rows);
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (s_scan[mid + 1] - j0 <= dl - mid - 1)
        lo = mid + 1;
      else
        hi = mid;
    }
    ri = lo;
    ej = dl - lo;
  }
  const int ri0 = ri, ej0 = ej;
  for (int s = 0; s < steps; ++s) {
    if (ri < n_rows && s_scan[ri + 1] - j0 <= ej) {
      ++ri;
    } else {
      s_row[ej] = (uint16_t)ri;
      ++ej;
    }
  }
  __syncthreads();

  // Phase 2: coalesced pass over the tile's edges: CSC source, gatherMap.
  // With delta_bits (delta-gather programs only), an edge whose source is not
  // in the bitmap contributes gatherZero instead of gatherMap: the row value
  // is already folded with gatherZero, and such a source cannot lower the
  // destination (it last gathered from the source after its last change), so
  // the result of apply is unchanged while the random source read is saved.
  {
    int src_v[VT], csc_v[VT];
#pragma unroll
    for (int m = 0; m < VT; ++m) {
      int e = tid + m * kNT;
      if (e < n_edges) {
        int k = s_row[e];
        csc_v[m] = s_base[k] + j0 + e;
        src_v[m] = __builtin_nontemporal_load(srcs + csc_v[m]);
      }
    }
    bool mapped = false;
    if constexpr (delta_gather_v<P>) {
      if (delta_bits) {
        mapped = true;
        uint32_t word[VT];
#pragma unroll
        for (int m = 0; m < VT; ++m) {
          int e = tid + m * kNT;
          word[m] = e < n_edges ? delta_bits[src_v[m] >> 5] : 0u;
        }
#pragma unroll
        for (int m = 0; m < VT; ++m) {
          int e = tid + m * kNT;
          if (e < n_edges) {
            if ((word[m] >> (src_v[m] & 31)) & 1u) {
              const typename P::EdgeData ev = stream_load(ed + edge_pos<P>(ed_index, csc_v[m]));
              s_val[e] = map_source<P>(vd, vd, srcval, src_v[m], &ev);
            } else {
              s_val[e] = P::gatherZero;
            }
          }
        }
      }
    }
    if (!mapped) {
#pragma unroll
      for (int m = 0; m < VT; ++m) {
        int e = tid + m * kNT;
        if (e < n_edges) {
          const typename P::VertexData* dstp = vd;
          if constexpr (P::kGatherUsesDst) dstp = vd + active[i0 + s_row[e]];
          const typename P::EdgeData ev = stream_load(ed + edge_pos<P>(ed_index, csc_v[m]));
          s_val[e] = map_source<P>(dstp, vd, srcval, src_v[m], &ev);
        }
      }
    }
  }
  __syncthreads();

  // Phase 3: sequential reduction along the lane's path items.
  ri = ri0;
  ej = ej0;
  G acc = G();
  int has = 0;
  int emitted = 0, first_row = 0, first_has = 0;
  G first_val = G();
  for (int s = 0; s < steps; ++s) {
    if (ri < n_rows && s_scan[ri + 1] - j0 <= ej) {
      if (!emitted) {
        emitted = 1;
        first_row = ri;
        first_val = acc;
        first_has = has;
      } else {
        gtmp[i0 + ri] = finalize<P>(has, acc);
      }
      has = 0;
      ++ri;
    } else {
      G v = s_val[ej];
      acc = has ? P::gatherReduce(acc, v) : v;
      has = 1;
      ++ej;
    }
  }

  // Segmented exclusive scan of the lanes' open tails.
  S mine;
  mine.flag = emitted;
  mine.has = has;
  mine.val = acc;
  const int lane = lane_id(), wv = tid / kWave;
  S incl = mine;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    S o = S::shfl_up(incl, off);
    if (lane >= off) incl = S::combine(o, incl);
  }
  if (lane == kWave - 1) s_wave[wv] = incl;
  __syncthreads();
  S wave_prefix = S::identity();
  for (int i = 0; i < wv; ++i) wave_prefix = S::combine(wave_prefix, s_wave[i]);
  S prev = S::shfl_up(incl, 1);
  S excl = (lane == 0) ? wave_prefix : S::combine(wave_prefix, prev);

  const bool head_open = n_rows > 0 && s_scan[0] < j0;  // row i0 began in an earlier tile
  if (emitted) {
    S fr;
    fr.flag = 0;
    fr.has = first_has;
    fr.val = first_val;
    S rv = S::combine(excl, fr);
    if (first_row == 0 && head_open) {
      Part<G> p;
      p.row = i0;
      p.has = rv.has;
      p.val = rv.val;
      head[t] = p;
    } else {
      gtmp[i0 + first_row] = finalize<P>(rv.has, rv.val);
    }
  }
  if (tid == 0 && !head_open) head[t].row = -1;
  if (tid == kNT - 1) {
    S all = S::identity();
    for (int i = 0; i < kNT / kWave; ++i) all = S::combine(all, s_wave[i]);
    Part<G> p;
    // The row open at the end carries over when some of its edges were here.
    if (i1 < n && escan[i1] < j1) {
      p.row = i1;
      p.has = all.has;
      p.val = all.val;
    } else {
      p.row = -1;
      p.has = 0;
      p.val = G();
    }
    carry[t] = p;
  }
}

// Rows spanning several tiles: carries in tile order, then the closing
// tile's head partial.
template <class P>
__global__ void k_gather_fixup(int n_tiles, const Part<typename P::GatherResult>* __restrict__ carry,
                               const Part<typename P::GatherResult>* __restrict__ head,
                               typename P::GatherResult* __restrict__ gtmp) {
  typedef typename P::GatherResult G;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  const int r = carry[t].row;
  if (r < 0 || (t > 0 && carry[t - 1].row == r)) return;
  int has = carry[t].has;
  G acc = carry[t].val;
  // A hub row spans many tiles: their carries are read kBatch at a time, so
  // that the walk costs one memory round trip per kBatch tiles, not per tile.
  // u ends at the first tile past t whose carry belongs to another row.
  constexpr int kBatch = 16;
  int u = t + 1;
  bool open = true;
  while (open && u < n_tiles) {
    const int cnt = min(kBatch, n_tiles - u);
    Part<G> c[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; ++k)
      if (k < cnt) c[k] = carry[u + k];
    int step = cnt;
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      if (open && k < cnt) {
        if (c[k].row != r) {
          open = false;
          step = k;
        } else if (c[k].has) {
          acc = has ? P::gatherReduce(acc, c[k].val) : c[k].val;
          has = 1;
        }
      }
    }
    u += step;
  }
  if (u < n_tiles && head[u].row == r && head[u].has) {
    acc = has ? P::gatherReduce(acc, head[u].val) : head[u].val;
    has = 1;
  }
  gtmp[r] = finalize<P>(has, acc);
}

// Direction rules of scatterActivate, shared by its kernels: pull when the
// changed out-edges (low half of chg) exceed the threshold (mode 0; 1 push
// only, 2 pull only); the fused activation + push gather of delta-gather
// programs when they number at most fuse_max (>= 0).
__device__ __forceinline__ bool use_pull_rule(unsigned long long chg, int mode,
                                              long long threshold) {
  if (mode == 2) return true;
  if (mode == 1) return false;
  return (long long)(uint32_t)chg > threshold;
}
__device__ __forceinline__ bool fused_scatter_rule(unsigned long long chg, long long fuse_max) {
  return fuse_max >= 0 && (chg >> 32) != 0 && (long long)(uint32_t)chg <= fuse_max;
}

// ------------------------------------------------------- blocked gather
//
// Gather of every row by row blocks (gr_blocked): one workgroup per block
// keeps the block's row values in LDS and streams the block's in-edges, which
// are listed by source, so that the random source reads of all workgroups
// advance through the vertex ids together and mostly hit the same cache
// lines.  For programs whose gatherReduce is an int minimum that gatherMap
// only feeds from the source and the edge (the delta-gather programs): the
// row value starts at gatherZero and falls by LDS atomicMin, which equals the
// reference's fold of gatherReduce from gatherZero over the in-edges
// (refgas.h:141-160) in any order.  delta_bits (optional) skips the sources
// outside the bitmap as k_gather_tiles does.  Writes gdense[row] for every row
// of every block (vertex-indexed).
//
// The in-edges come packed (gr_blocked): each wave takes one group of
// kPackGroup slots per pass, lane l its slots l, l + 64, ... (one coalesced
// 4-byte load per 64 slots, so that the 64 sources of one load are
// consecutive in the block's source order and their random value reads share
// cache lines); the sources are the group's base plus the prefix sums of the
// steps, a wave scan per 64 slots.
constexpr int kBlkNT = 1024;  // one workgroup per CU (128 KB of LDS): 16 waves
constexpr int kBlkVT = 8;
static_assert(kBlkVT * kWave == kPackGroup, "a wave's pass is one group");

// This lane's packed slots of group `gs` (its first slot).
__device__ __forceinline__ void load_slot_words(const uint32_t* __restrict__ bpk, int gs,
                                                uint32_t (&w)[kBlkVT]) {
  const int lane = lane_id();
#pragma unroll
  for (int m = 0; m < kBlkVT; ++m) w[m] = __builtin_nontemporal_load(bpk + gs + m * kWave + lane);
}

// The sources and rows of this lane's slots of a group whose first slot's
// source is `base`; a pad slot yields source -1.  Every lane of the wave
// calls it together.
__device__ __forceinline__ void decode_slots(const uint32_t (&w)[kBlkVT], int base,
                                             int (&src)[kBlkVT], int (&row)[kBlkVT]) {
  int carry = base;
#pragma unroll
  for (int m = 0; m < kBlkVT; ++m) {
    const int incl = wave_incl_scan_add_dpp((int)(w[m] >> 16));
    row[m] = (int)(w[m] & 0xFFFFu);
    src[m] = row[m] == (int)kPackPad ? -1 : carry + incl;
    carry += lane_value(incl, kWave - 1);
  }
}

// Edge data of this lane's slots of group `gs` where src >= 0.
template <class ED>
__device__ __forceinline__ void load_slot_edges(const ED* __restrict__ ed, int gs,
                                                const int (&src)[kBlkVT], ED (&ev)[kBlkVT]) {
  const int lane = lane_id();
#pragma unroll
  for (int m = 0; m < kBlkVT; ++m)
    if (src[m] >= 0) ev[m] = stream_load(ed + gs + m * kWave + lane);
}

//
// Fused activation (frontier != nullptr): the kernel is launched by
// scatterActivate in place of the pull and runs only when the pull would
// (`run` decides from the counters, the same for every workgroup; workgroup 0
// records the outcome in ctr->fab).  delta_bits is then the bitmap of the
// vertices whose apply just returned true: the values are the next delta
// gather's, and a row with an in-edge from such a vertex is marked in
// `frontier` (which holds no bit of these rows before), as the pull marks it.
// kRows is the blocks' row limit (the LDS minima): kBlockRows, one workgroup
// per CU, or kBlockRowsWide, two (40-50 VGPRs leave room for 8 waves per SIMD).
template <class P, int kRows = kMinBlockRows>
__global__ __launch_bounds__(kBlkNT) void k_gather_blocked(
    const int32_t* __restrict__ brow, const int32_t* __restrict__ bnr,
    const int32_t* __restrict__ boff, const uint32_t* __restrict__ bpk,
    const int32_t* __restrict__ gbase, const typename P::EdgeData* __restrict__ ed,
    const typename P::VertexData* __restrict__ vd, const source_value_t<P>* __restrict__ srcval,
    const uint32_t* __restrict__ delta_bits, typename P::GatherResult* __restrict__ gdense,
    uint32_t* __restrict__ frontier = nullptr, DevCounters* __restrict__ ctr = nullptr,
    int mode = 0, long long threshold = 0, long long fuse_max = -1) {
  typedef typename P::GatherResult G;
  static_assert(std::is_same<G, int>::value && delta_gather_v<P> && !P::kGatherUsesDst,
                "the blocked gather folds int minima of source-only gatherMaps");
  __shared__ int acc[kRows];
  __shared__ uint32_t act[kRows / 32];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  if (frontier) {
    const unsigned long long c = ctr->chg;
    const bool run = use_pull_rule(c, mode, threshold) && !fused_scatter_rule(c, fuse_max);
    if (b == 0 && tid == 0) ctr->fab = run ? 1 : 0;
    if (!run) return;
  }
  const int r0 = brow[b], nr = bnr[b];
  const int e0 = boff[b], e1 = boff[b + 1];
  for (int r = tid; r < nr; r += kBlkNT) acc[r] = P::gatherZero;
  for (int k = tid; k < kRows / 32; k += kBlkNT) act[k] = 0u;
  __syncthreads();
  // each wave's next group is loaded while it works on the current one
  constexpr int kPass = kBlkNT * kBlkVT;
  const int wv = __builtin_amdgcn_readfirstlane(tid / kWave);
  uint32_t wcur[kBlkVT] = {};
  if (e0 + wv * kPackGroup < e1) load_slot_words(bpk, e0 + wv * kPackGroup, wcur);
  for (int gs = e0 + wv * kPackGroup; gs < e1; gs += kPass) {
    uint32_t wnext[kBlkVT] = {};
    if (gs + kPass < e1) load_slot_words(bpk, gs + kPass, wnext);
    int src[kBlkVT], row[kBlkVT];
    typename P::EdgeData ev[kBlkVT];
    decode_slots(wcur, gbase[gs / kPackGroup], src, row);
#pragma unroll
    for (int m = 0; m < kBlkVT; ++m) wcur[m] = wnext[m];
    if (delta_bits) {
      uint32_t word[kBlkVT];
#pragma unroll
      for (int m = 0; m < kBlkVT; ++m) word[m] = src[m] >= 0 ? delta_bits[src[m] >> 5] : 0u;
#pragma unroll
      for (int m = 0; m < kBlkVT; ++m)
        if (!((word[m] >> (src[m] & 31)) & 1u)) src[m] = -1;
    }
    // edge data only for the edges that count
    load_slot_edges(ed, gs, src, ev);
    G val[kBlkVT];
#pragma unroll
    for (int m = 0; m < kBlkVT; ++m)
      val[m] = src[m] >= 0 ? map_source<P>(vd, vd, srcval, src[m], &ev[m]) : P::gatherZero;
    // a value not below gatherZero cannot lower a row that starts there
    // (activation: a row whose minimum ends below gatherZero had an in-edge
    // from a changed source; only the other such edges need a mark)
#pragma unroll
    for (int m = 0; m < kBlkVT; ++m) {
      if (src[m] >= 0 && val[m] < P::gatherZero)
        atomicMin(&acc[row[m]], val[m]);
      else if (frontier && src[m] >= 0)
        atomicOr(&act[row[m] >> 5], 1u << (row[m] & 31));
    }
  }
  __syncthreads();
  for (int r = tid; r < nr; r += kBlkNT) gdense[r0 + r] = acc[r];
  if (frontier) {
    // the bitmap words of rows r0 .. r0+nr-1; the first and last may hold
    // rows of the neighbouring blocks, hence atomicOr
    const int w0 = r0 >> 5, w1 = (r0 + nr - 1) >> 5;
    for (int w = w0 + tid; w <= w1; w += kBlkNT) {
      uint32_t bits = 0u;
      for (int k = 0; k < 32; ++k) {
        const int r = w * 32 + k - r0;
        if (r >= 0 && r < nr &&
            (acc[r] < P::gatherZero || ((act[r >> 5] >> (r & 31)) & 1u)))
          bits |= 1u << k;
      }
      if (bits) atomicOr(&frontier[w], bits);
    }
  }
}

constexpr float kFixScale = 4294967296.0f;             // 2^32
constexpr double kFixUnscale = 1.0 / 4294967296.0;     // 2^-32

// Blocked gather of a fixed-point sum program: as k_gather_blocked, with the
// rows' sums in 64-bit integers (kBlockRowsWide rows of 8 bytes in LDS).
// With chg_bits (the vertices whose apply just returned true), the same pass
// is also scatterActivate's pull: a row with an in-edge from such a vertex is
// marked in the `frontier` bitmap (which holds no bit of these rows before).
// It then stands down when ctr->chg shows no changed vertex with out-edges:
// the next active set is empty and no gather follows.
template <class P>
__global__ __launch_bounds__(kBlkNT) void k_gather_blocked_sum(
    const int32_t* __restrict__ brow, const int32_t* __restrict__ bnr,
    const int32_t* __restrict__ boff, const uint32_t* __restrict__ bpk,
    const int32_t* __restrict__ gbase, const typename P::EdgeData* __restrict__ ed,
    const typename P::VertexData* __restrict__ vd, const source_value_t<P>* __restrict__ srcval,
    typename P::GatherResult* __restrict__ gdense, const uint32_t* __restrict__ chg_bits,
    uint32_t* __restrict__ frontier, const DevCounters* __restrict__ ctr, int use_flags = 0) {
  typedef typename P::GatherResult G;
  static_assert(std::is_same<G, float>::value && fixed_point_sum_v<P> && !P::kGatherUsesDst,
                "the fixed-point blocked gather sums non-negative floats");
  __shared__ unsigned long long acc[kBlockRowsWide];
  __shared__ uint8_t act[kBlockRowsWide];
  if (frontier && ctr->chg == 0) return;  // uniform: nothing to activate or gather
  const int b = blockIdx.x;
  const int r0 = brow[b], nr = bnr[b];
  const int e0 = boff[b], e1 = boff[b + 1];
  const int tid = threadIdx.x;
  for (int r = tid; r < nr; r += kBlkNT) {
    acc[r] = 0ull;
    act[r] = 0;
  }
  __syncthreads();
  constexpr int kPass = kBlkNT * kBlkVT;
  const int wv = __builtin_amdgcn_readfirstlane(tid / kWave);
  uint32_t wcur[kBlkVT] = {};
  if (e0 + wv * kPackGroup < e1) load_slot_words(bpk, e0 + wv * kPackGroup, wcur);
  for (int gs = e0 + wv * kPackGroup; gs < e1; gs += kPass) {
    uint32_t wnext[kBlkVT] = {};
    if (gs + kPass < e1) load_slot_words(bpk, gs + kPass, wnext);
    int src[kBlkVT], row[kBlkVT];
    typename P::EdgeData ev[kBlkVT];
    decode_slots(wcur, gbase[gs / kPackGroup], src, row);
#pragma unroll
    for (int m = 0; m < kBlkVT; ++m) wcur[m] = wnext[m];
    load_slot_edges(ed, gs, src, ev);
    unsigned long long fix[kBlkVT];
    bool chg[kBlkVT];
    if constexpr (source_flag_v<P>) {
      // every lane's loads first (all in flight together), then the terms
      uint32_t raw[kBlkVT], word[kBlkVT];
#pragma unroll
      for (int m = 0; m < kBlkVT; ++m) raw[m] = src[m] >= 0 ? __float_as_uint(srcval[src[m]]) : 0u;
      const bool bits = chg_bits && !use_flags;  // uniform
#pragma unroll
      for (int m = 0; m < kBlkVT; ++m) word[m] = bits && src[m] >= 0 ? chg_bits[src[m] >> 5] : 0u;
#pragma unroll
      for (int m = 0; m < kBlkVT; ++m) {
        const G v = P::gatherMapValue(__uint_as_float(raw[m] & 0x7FFFFFFFu), &ev[m]);
        fix[m] = src[m] >= 0 ? (unsigned long long)__float2ll_rn(__fmul_rn(v, kFixScale)) : 0ull;
        chg[m] = src[m] >= 0 && chg_bits &&
                 (use_flags ? (raw[m] >> 31) != 0u : ((word[m] >> (src[m] & 31)) & 1u) != 0u);
      }
    } else {
      (void)use_flags;
      uint32_t word[kBlkVT];
#pragma unroll
      for (int m = 0; m < kBlkVT; ++m) {
        fix[m] = 0ull;
        word[m] = 0u;
        if (src[m] >= 0) {
          const G v = map_source<P>(vd, vd, srcval, src[m], &ev[m]);
          fix[m] = (unsigned long long)__float2ll_rn(__fmul_rn(v, kFixScale));
          if (chg_bits) word[m] = chg_bits[src[m] >> 5];
        }
      }
#pragma unroll
      for (int m = 0; m < kBlkVT; ++m)
        chg[m] = src[m] >= 0 && ((word[m] >> (src[m] & 31)) & 1u);
    }
#pragma unroll
    for (int m = 0; m < kBlkVT; ++m) {
      if (fix[m]) atomicAdd(&acc[row[m]], fix[m]);
      if (chg[m]) act[row[m]] = 1;
    }
  }
  __syncthreads();
  for (int r = tid; r < nr; r += kBlkNT)
    gdense[r0 + r] = (G)((double)(long long)acc[r] * kFixUnscale);
  if (chg_bits) {
    // the bitmap words of rows r0 .. r0+nr-1; the first and last may hold
    // rows of the neighbouring blocks, hence atomicOr
    const int w0 = r0 >> 5, w1 = (r0 + nr - 1) >> 5;
    for (int w = w0 + tid; w <= w1; w += kBlkNT) {
      uint32_t bits = 0u;
      for (int k = 0; k < 32; ++k) {
        const int r = w * 32 + k - r0;
        if (r >= 0 && r < nr && act[r]) bits |= 1u << k;
      }
      if (bits) atomicOr(&frontier[w], bits);
    }
  }
}

}  // namespace gr

#endif
