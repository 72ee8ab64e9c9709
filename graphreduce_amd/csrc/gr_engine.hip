// Gather-Apply-Scatter engine for one MI355X: the whole CSR + CSC graph and
// all engine state live in HBM; every phase is a hand-written HIP kernel.
//
// One iteration (refgas.h:220-229 semantics, bulk-synchronous):
//   gather           k_merge_partition + k_gather_tiles + k_gather_fixup
//                    merge-path split of the active vertices' in-edge ranges
//                    into fixed tiles of kTile path items, coalesced CSC
//                    reads staged in LDS, per-lane sequential reduction and a
//                    wave64 segmented scan across lanes; rows that span tiles
//                    are combined in tile order by the fix-up kernel.
//   apply            k_apply: Program::apply per active vertex; vertices whose
//                    apply returned true are appended (wave-aggregated atomic)
//                    to the changed list together with their out-edge offsets.
//   scatterActivate  push: k_push_partition + k_push_tiles set one bit per
//                    out-neighbour in the next-frontier bitmap; pull:
//                    k_pull_activate tests each vertex's in-edges against the
//                    changed bitmap.
//   nextIter         k_compact_count/scan/write: bitmap -> ascending active
//                    list plus the exclusive scan of its in-degrees (the
//                    gather's row offsets); one 32-byte device->host read
//                    returns the new counts.
#include <string.h>

#include <type_traits>
#include <vector>

#include "gr_device.h"
#include "gr_programs.h"

namespace gr {

constexpr int kNT = 256;            // threads per tile workgroup (4 waves)
constexpr int kVT = 8;              // merge-path items per thread
constexpr int kTile = kNT * kVT;    // merge-path items per tile
constexpr int kCompactNT = 256;     // compaction: one bitmap word per thread
constexpr int kMaxGrid = 1 << 20;

struct DevCounters {
  int32_t n_active;               // size of the active list built by nextIter
  int32_t w_edges;                // sum of in-degrees over that list
  unsigned long long chg;         // (#changed << 32) | out-edges, appended by apply
  unsigned long long chg_last;    // chg as seen by the last nextIter
  int32_t pull;                   // last scatterActivate ran in pull mode
  int32_t pad[3];
};

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

__global__ void k_merge_partition(const int32_t* __restrict__ escan, int n, int w, int n_tiles,
                                  int32_t* __restrict__ part) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  int64_t d = min((int64_t)t * kTile, (int64_t)n + w);
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
    r.val = __shfl_up(s.val, off, kWave);
    return r;
  }
};

template <class P>
__device__ __forceinline__ typename P::GatherResult finalize(int has,
                                                             typename P::GatherResult v) {
  return has ? P::gatherReduce(P::gatherZero, v) : P::gatherZero;
}

// ------------------------------------------------------------------ gather

template <class P>
__global__ __launch_bounds__(kNT) void k_gather_tiles(
    int n, int w, const int32_t* __restrict__ active, const int32_t* __restrict__ escan,
    const int32_t* __restrict__ part, const int32_t* __restrict__ off_in,
    const int32_t* __restrict__ srcs, const typename P::VertexData* __restrict__ vd,
    const typename P::EdgeData* __restrict__ ed, typename P::GatherResult* __restrict__ gtmp,
    Part<typename P::GatherResult>* __restrict__ carry,
    Part<typename P::GatherResult>* __restrict__ head) {
  typedef typename P::GatherResult G;
  typedef Seg<P> S;
  __shared__ int32_t s_scan[kTile + 2];
  __shared__ int32_t s_base[kTile + 1];
  __shared__ uint16_t s_row[kTile];
  __shared__ G s_val[kTile];
  __shared__ S s_wave[kNT / kWave];

  const int t = blockIdx.x;
  const int tid = threadIdx.x;
  const int total = n + w;
  const int d0 = t * kTile;
  const int d1 = min(d0 + kTile, total);
  const int i0 = part[t], i1 = part[t + 1];
  const int j0 = d0 - i0, j1 = d1 - i1;
  const int n_rows = i1 - i0;                    // row ends inside this tile
  const int n_ref = n_rows + (i1 < n ? 1 : 0);   // plus the row still open at the end
  const int n_edges = j1 - j0;

  for (int k = tid; k <= n_rows; k += kNT) s_scan[k] = escan[i0 + k];
  for (int k = tid; k < n_ref; k += kNT) s_base[k] = off_in[active[i0 + k]] - escan[i0 + k];
  __syncthreads();

  // Phase 1: each lane walks its kVT path items and labels edges with rows.
  const int dl = tid * kVT;
  const int steps = max(0, min(kVT, (d1 - d0) - dl));
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
  {
    int src_v[kVT], csc_v[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      int e = tid + m * kNT;
      if (e < n_edges) {
        int k = s_row[e];
        csc_v[m] = s_base[k] + j0 + e;
        src_v[m] = srcs[csc_v[m]];
      }
    }
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      int e = tid + m * kNT;
      if (e < n_edges) {
        const typename P::VertexData* dstp = vd;
        if constexpr (P::kGatherUsesDst) dstp = vd + active[i0 + s_row[e]];
        s_val[e] = P::gatherMap(dstp, vd + src_v[m], ed + csc_v[m]);
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
  int u = t + 1;
  while (u < n_tiles && carry[u].row == r) {
    if (carry[u].has) {
      acc = has ? P::gatherReduce(acc, carry[u].val) : carry[u].val;
      has = 1;
    }
    ++u;
  }
  if (u < n_tiles && head[u].row == r && head[u].has) {
    acc = has ? P::gatherReduce(acc, head[u].val) : head[u].val;
    has = 1;
  }
  gtmp[r] = finalize<P>(has, acc);
}

// ------------------------------------------------------------------- apply

// Appends the lanes with want != 0 to the changed list, keeping the
// out-edge offsets increasing with the list index.
__device__ __forceinline__ void append_changed(bool want, int v, int deg,
                                               int32_t* __restrict__ chg_v,
                                               int32_t* __restrict__ chg_e,
                                               unsigned long long* __restrict__ counter) {
  const uint64_t mask = __ballot(want);
  if (mask == 0) return;
  const int lane = lane_id();
  const int e = want ? deg : 0;
  const int incl = wave_incl_scan_add(e);
  const int total = __shfl(incl, kWave - 1, kWave);
  const int cnt = __popcll(mask);
  unsigned long long base = 0;
  if (lane == 0)
    base = atomicAdd(counter, ((unsigned long long)cnt << 32) | (unsigned long long)(uint32_t)total);
  base = __shfl(base, 0, kWave);
  if (want) {
    const int idx = (int)(base >> 32) + lanes_below(mask);
    chg_v[idx] = v;
    chg_e[idx] = (int)(uint32_t)base + incl - e;
  }
}

template <class T>
__device__ __forceinline__ bool pod_equal(const T& a, const T& b) {
  static_assert(sizeof(T) % 4 == 0, "vertex data must be a multiple of 4 bytes");
  const int* pa = reinterpret_cast<const int*>(&a);
  const int* pb = reinterpret_cast<const int*>(&b);
  bool eq = true;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) eq = eq && (pa[i] == pb[i]);
  return eq;
}

template <class P>
__global__ __launch_bounds__(256) void k_apply(
    int n, const int32_t* __restrict__ active, const typename P::GatherResult* __restrict__ gtmp,
    int use_gather, typename P::VertexData* __restrict__ vd, const int32_t* __restrict__ off_out,
    int iteration, int32_t* __restrict__ chg_v, int32_t* __restrict__ chg_e,
    uint32_t* __restrict__ chg_bits, DevCounters* __restrict__ ctr) {
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int nw = gridDim.x * wpb;
  for (int base = (blockIdx.x * wpb + threadIdx.x / kWave) * kWave; base < n; base += nw * kWave) {
    const int i = base + lane;
    bool want = false;
    int dv = 0, deg = 0;
    if (i < n) {
      dv = active[i];
      typename P::GatherResult g = use_gather ? gtmp[i] : P::gatherZero;
      typename P::VertexData v = vd[dv];
      const typename P::VertexData old = v;
      bool r = P::apply(&v, g, iteration);
      if (!pod_equal(v, old)) vd[dv] = v;
      if (r) {
        deg = off_out[dv + 1] - off_out[dv];
        want = deg > 0;
        if (want && chg_bits) atomicOr(&chg_bits[dv >> 5], 1u << (dv & 31));
      }
    }
    append_changed(want, dv, deg, chg_v, chg_e, &ctr->chg);
  }
}

// ------------------------------------------------------- scatter: push

// Direction of scatterActivate, decided on the device from the changed
// out-edge count so that the host needs no read: mode 0 = pull when the
// out-edges of the changed vertices exceed `threshold`, 1 = push, 2 = pull.
__device__ __forceinline__ bool use_pull(unsigned long long chg, int mode, long long threshold) {
  if (mode == 2) return true;
  if (mode == 1) return false;
  return (long long)(uint32_t)chg > threshold;
}

__device__ __forceinline__ int upper_bound_dev(const int32_t* __restrict__ a, int n, int x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// part[t] = changed-list entry holding out-edge t * kTile (t <= n_tiles).
__global__ void k_push_partition(const DevCounters* __restrict__ ctr, int mode,
                                 long long threshold, const int32_t* __restrict__ chg_e,
                                 int32_t* __restrict__ part, int max_tiles) {
  const unsigned long long c = ctr->chg;
  if (use_pull(c, mode, threshold)) return;
  const int n_c = (int)(c >> 32);
  const int w_c = (int)(uint32_t)c;
  const int n_tiles = (w_c + kTile - 1) / kTile;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= n_tiles && t <= max_tiles;
       t += gridDim.x * blockDim.x) {
    int e = min(t * kTile, w_c);
    part[t] = max(0, upper_bound_dev(chg_e, n_c, e) - 1);
  }
}

__global__ __launch_bounds__(kNT) void k_push_tiles(const DevCounters* __restrict__ ctr,
                                                    int mode, long long threshold,
                                                    const int32_t* __restrict__ part,
                                                    const int32_t* __restrict__ chg_v,
                                                    const int32_t* __restrict__ chg_e,
                                                    const int32_t* __restrict__ off_out,
                                                    const int32_t* __restrict__ dsts,
                                                    uint32_t* __restrict__ bitmap) {
  __shared__ int32_t s_start[kTile + 2];
  __shared__ int32_t s_csr[kTile + 2];
  const unsigned long long c = ctr->chg;
  if (use_pull(c, mode, threshold)) return;
  const int w_c = (int)(uint32_t)c;
  const int n_tiles = (w_c + kTile - 1) / kTile;
  const int tid = threadIdx.x;
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int k0 = part[t], k1 = part[t + 1];
    const int nr = k1 - k0 + 1;
    const int e0 = t * kTile, e1 = min(e0 + kTile, w_c);
    for (int k = tid; k < nr; k += kNT) {
      s_start[k] = chg_e[k0 + k];
      s_csr[k] = off_out[chg_v[k0 + k]];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      const int e = e0 + tid + m * kNT;
      if (e < e1) {
        int lo = 0, hi = nr - 1;
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (s_start[mid] <= e)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int dst = dsts[s_csr[lo] + (e - s_start[lo])];
        atomicOr(&bitmap[dst >> 5], 1u << (dst & 31));
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------- scatter: pull

// next[v] = some in-neighbour of v is in the changed bitmap.  One wave owns
// 64 consecutive vertices: light vertices are scanned by their own lane,
// heavy ones (in-degree > kPullLight) by the whole wave with early exit.
constexpr int kPullLight = 32;

__global__ __launch_bounds__(256) void k_pull_activate(
    const DevCounters* __restrict__ ctr, int mode, long long threshold, int nV,
    const int32_t* __restrict__ off_in, const int32_t* __restrict__ srcs,
    const uint32_t* __restrict__ chg_bits, uint32_t* __restrict__ bitmap) {
  if (!use_pull(ctr->chg, mode, threshold)) return;
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int nw = gridDim.x * wpb;
  for (int base = (blockIdx.x * wpb + threadIdx.x / kWave) * kWave; base < nV; base += nw * kWave) {
    const int v = base + lane;
    int b = 0, e = 0;
    if (v < nV) {
      b = off_in[v];
      e = off_in[v + 1];
    }
    const bool light = (e - b) <= kPullLight;
    bool hit = false;
    if (v < nV && light) {
      for (int p = b; p < e; ++p) {
        const int u = srcs[p];
        if ((chg_bits[u >> 5] >> (u & 31)) & 1u) {
          hit = true;
          break;
        }
      }
    }
    uint64_t heavy = __ballot(v < nV && !light);
    while (heavy) {
      const int hl = __builtin_ctzll(heavy);
      heavy &= heavy - 1;
      const int hb = __shfl(b, hl, kWave), he = __shfl(e, hl, kWave);
      // wave-uniform trip count, so every lane sees the same `found`
      bool found = false;
      for (int pb = hb; pb < he; pb += kWave) {
        const int p = pb + lane;
        bool h = false;
        if (p < he) {
          const int u = srcs[p];
          h = (chg_bits[u >> 5] >> (u & 31)) & 1u;
        }
        if (__ballot(h)) {
          found = true;
          break;
        }
      }
      if (lane == hl) hit = found;
    }
    const uint64_t hits = __ballot(hit);
    // lanes 0 and 32 write the two 32-bit words of this wave's 64 vertices
    if (v < nV && (lane & 31) == 0) {
      const uint32_t word = (uint32_t)(hits >> lane);
      if (word) bitmap[v >> 5] = word;
    }
  }
}

// -------------------------------------------------------------- compaction

__device__ __forceinline__ int degree_of_bits(uint32_t word, int base_v,
                                              const int32_t* __restrict__ off) {
  int deg = 0;
  while (word) {
    const int b = __builtin_ctz(word);
    word &= word - 1;
    const int v = base_v + b;
    deg += off[v + 1] - off[v];
  }
  return deg;
}

__global__ __launch_bounds__(kCompactNT) void k_compact_count(const uint32_t* __restrict__ bitmap,
                                                              int n_words,
                                                              const int32_t* __restrict__ off_in,
                                                              int need_deg, int2* __restrict__ blk) {
  __shared__ int s_a[kCompactNT / kWave], s_b[kCompactNT / kWave];
  const int w = blockIdx.x * kCompactNT + threadIdx.x;
  const uint32_t word = (w < n_words) ? bitmap[w] : 0u;
  int cnt = __popc(word);
  int deg = need_deg ? degree_of_bits(word, w * 32, off_in) : 0;
  cnt = wave_reduce_add(cnt);
  deg = wave_reduce_add(deg);
  if (lane_id() == 0) {
    s_a[threadIdx.x / kWave] = cnt;
    s_b[threadIdx.x / kWave] = deg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int i = 0; i < kCompactNT / kWave; ++i) {
      a += s_a[i];
      b += s_b[i];
    }
    blk[blockIdx.x] = make_int2(a, b);
  }
}

// Single workgroup: exclusive scan of the per-block counts; publishes the
// totals and resets the changed-list counter for the next apply.
__global__ __launch_bounds__(1024) void k_compact_scan(int2* __restrict__ blk, int nb,
                                                       DevCounters* __restrict__ ctr,
                                                       int32_t* __restrict__ escan, int mode,
                                                       long long threshold) {
  __shared__ int s_a[1024 / kWave], s_b[1024 / kWave];
  int run_a = 0, run_b = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int2 v = (i < nb) ? blk[i] : make_int2(0, 0);
    int ta, tb;
    const int ea = block_excl_scan_add<1024>(v.x, s_a, &ta);
    const int eb = block_excl_scan_add<1024>(v.y, s_b, &tb);
    if (i < nb) blk[i] = make_int2(run_a + ea, run_b + eb);
    run_a += ta;
    run_b += tb;
  }
  if (threadIdx.x == 0) {
    ctr->n_active = run_a;
    ctr->w_edges = run_b;
    if (escan) escan[run_a] = run_b;
    ctr->chg_last = ctr->chg;
    ctr->pull = use_pull(ctr->chg, mode, threshold) ? 1 : 0;
    ctr->chg = 0;
  }
}

__global__ __launch_bounds__(kCompactNT) void k_compact_write(
    uint32_t* __restrict__ bitmap, int n_words, const int32_t* __restrict__ off_in, int need_deg,
    const int2* __restrict__ blk, int32_t* __restrict__ active, int32_t* __restrict__ escan,
    uint32_t* __restrict__ chg_bits) {
  __shared__ int s_a[kCompactNT / kWave], s_b[kCompactNT / kWave];
  const int w = blockIdx.x * kCompactNT + threadIdx.x;
  // the changed bitmap (pull-mode input) starts empty for the next apply
  if (chg_bits && w < n_words) chg_bits[w] = 0u;
  uint32_t word = (w < n_words) ? bitmap[w] : 0u;
  const int cnt = __popc(word);
  const int deg = need_deg ? degree_of_bits(word, w * 32, off_in) : 0;
  int ta, tb;
  const int ea = block_excl_scan_add<kCompactNT>(cnt, s_a, &ta);
  const int eb = block_excl_scan_add<kCompactNT>(deg, s_b, &tb);
  if (!word) return;
  const int2 b = blk[blockIdx.x];
  int pos = b.x + ea, dpos = b.y + eb;
  bitmap[w] = 0u;
  while (word) {
    const int bit = __builtin_ctz(word);
    word &= word - 1;
    const int v = w * 32 + bit;
    active[pos] = v;
    if (need_deg) {
      escan[pos] = dpos;
      dpos += off_in[v + 1] - off_in[v];
    }
    ++pos;
  }
}

__global__ void k_set_active_range(int start, int n, const int32_t* __restrict__ off_in,
                                   int32_t* __restrict__ active, int32_t* __restrict__ escan,
                                   DevCounters* __restrict__ ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int base = off_in[start];
  if (i < n) active[i] = start + i;
  escan[i] = off_in[start + i] - base;
  if (i == n) {
    ctr->n_active = n;
    ctr->w_edges = off_in[start + n] - base;
    ctr->chg = 0;
  }
}

// Edge data from edge-list order into CSC order, in 4-byte words.
__global__ void k_permute_words(const uint32_t* __restrict__ in, const int32_t* __restrict__ eid,
                                int64_t n, int words, uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = eid ? (int64_t)eid[i] : i;
  for (int k = 0; k < words; ++k) out[i * words + k] = in[s * words + k];
}

// ================================================================ host side

struct EngineBase {
  virtual ~EngineBase() {}
  virtual int bind(gr_graph* g, bool owns, void* vdata, size_t velem, const void* edata,
                   size_t eelem, bool edata_on_device) = 0;
  virtual int set_active(int32_t start, int32_t end) = 0;
  virtual int gather() = 0;
  virtual int apply() = 0;
  virtual int scatter_activate(bool have_scatter) = 0;
  virtual int next_iter() = 0;
  virtual int get_results() = 0;
  virtual int read_vertex_data(void* out, size_t velem) = 0;
  virtual int reset_vertex_data() = 0;
  virtual int sync_stats(gr_engine_stats* out) = 0;
  virtual size_t vertex_size() const = 0;
  virtual size_t edge_size() const = 0;  // 0 for an empty EdgeData

  int32_t n_active = 0;
  int32_t w_edges = 0;
  int32_t iteration = 0;
  bool bound = false;
  int dir_mode = 0;
  double dir_alpha = 15.0;
  bool timing = false;
  gr_engine_stats stats{};
};

template <class P>
struct Engine final : EngineBase {
  typedef typename P::VertexData VD;
  typedef typename P::EdgeData ED;
  typedef typename P::GatherResult G;
  static constexpr bool kEdgeEmpty = std::is_empty<ED>::value;

  hipStream_t stream = nullptr;
  gr_graph* g = nullptr;
  bool owns_graph = false;
  void* host_vdata = nullptr;
  VD* d_vd = nullptr;
  VD* d_vd_init = nullptr;  // vertex data as given at bind time
  ED* d_ed = nullptr;  // CSC order
  int32_t* d_active = nullptr;
  int32_t* d_escan = nullptr;
  G* d_gtmp = nullptr;
  int32_t* d_chg_v = nullptr;
  int32_t* d_chg_e = nullptr;
  uint32_t* d_bitmap = nullptr;
  uint32_t* d_chg_bits = nullptr;
  int2* d_blk = nullptr;
  int32_t* d_part = nullptr;
  int32_t* d_push_part = nullptr;
  Part<G>* d_carry = nullptr;
  Part<G>* d_head = nullptr;
  DevCounters* d_ctr = nullptr;
  DevCounters* h_ctr = nullptr;  // pinned
  int n_words = 0, n_cblocks = 0, max_tiles = 0, max_push_tiles = 0;
  // hipEvent start/stop pairs per timed phase
  enum { kPhGather = 0, kPhScatter, kPhApply, kPhNext, kPhases };
  std::vector<hipEvent_t> ev[kPhases];
  hipEvent_t ev_open = nullptr;

  ~Engine() override { release(); }

  size_t vertex_size() const override { return sizeof(VD); }
  size_t edge_size() const override { return kEdgeEmpty ? 0 : sizeof(ED); }

  void release() {
    void* ptrs[] = {d_vd,    d_vd_init, d_ed,      d_active,    d_escan, d_gtmp,  d_chg_v,
                    d_chg_e, d_bitmap,  d_chg_bits,  d_blk,   d_part,  d_push_part,
                    d_carry, d_head,    d_ctr};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (h_ctr) (void)hipHostFree(h_ctr);
    for (auto& v : ev) {
      for (hipEvent_t e : v) (void)hipEventDestroy(e);
      v.clear();
    }
    if (stream) (void)hipStreamDestroy(stream);
    if (owns_graph && g) graph_free(g);
    d_vd = nullptr;
    d_vd_init = nullptr;
    d_ed = nullptr;
    d_active = d_escan = d_chg_v = d_chg_e = d_part = d_push_part = nullptr;
    d_gtmp = nullptr;
    d_bitmap = d_chg_bits = nullptr;
    d_blk = nullptr;
    d_carry = d_head = nullptr;
    d_ctr = h_ctr = nullptr;
    stream = nullptr;
    g = nullptr;
    bound = false;
  }

  template <class T>
  int dalloc(T** p, int64_t n) {
    if (n < 1) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
      *p = nullptr;
      return set_errorf(GR_ERR_NOMEM, "engine: hipMalloc(%lld bytes): %s",
                        (long long)(sizeof(T) * n), hipGetErrorString(e));
    }
    return GR_OK;
  }

  int bind(gr_graph* graph, bool owns, void* vdata, size_t velem, const void* edata,
           size_t eelem, bool edata_on_device) override {
    if (velem != sizeof(VD))
      return set_errorf(GR_ERR_INVALID, "vertex element size %zu != program VertexData size %zu",
                        velem, sizeof(VD));
    if (!kEdgeEmpty && edata && eelem != sizeof(ED))
      return set_errorf(GR_ERR_INVALID, "edge element size %zu != program EdgeData size %zu",
                        eelem, sizeof(ED));
    if (!kEdgeEmpty && !edata)
      return set_error(GR_ERR_INVALID, "this program needs edge data");
    if (!vdata) return set_error(GR_ERR_INVALID, "vertex data is NULL");
    release();
    g = graph;
    owns_graph = owns;
    host_vdata = vdata;
    const int64_t V = g->nV, E = g->nE;
    if (V + E >= ((int64_t)1 << 31) - kTile)
      return set_error(GR_ERR_INVALID, "graph too large for int32 merge-path coordinates");
    GR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    n_words = (int)((V + 31) / 32);
    n_cblocks = (n_words + kCompactNT - 1) / kCompactNT;
    max_tiles = (int)((V + E + kTile - 1) / kTile) + 1;
    max_push_tiles = (int)((E + kTile - 1) / kTile) + 1;
    GR_TRY(dalloc(&d_vd, V));
    GR_TRY(dalloc(&d_vd_init, V));
    GR_TRY(dalloc(&d_active, V));
    GR_TRY(dalloc(&d_escan, V + 1));
    GR_TRY(dalloc(&d_gtmp, V));
    GR_TRY(dalloc(&d_chg_v, V));
    GR_TRY(dalloc(&d_chg_e, V));
    GR_TRY(dalloc(&d_bitmap, n_words + 1));
    GR_TRY(dalloc(&d_chg_bits, n_words + 1));
    GR_TRY(dalloc(&d_blk, n_cblocks));
    GR_TRY(dalloc(&d_part, max_tiles + 2));
    GR_TRY(dalloc(&d_push_part, max_push_tiles + 2));
    GR_TRY(dalloc(&d_carry, max_tiles + 1));
    GR_TRY(dalloc(&d_head, max_tiles + 1));
    GR_TRY(dalloc(&d_ctr, 1));
    GR_HIP(hipHostMalloc((void**)&h_ctr, sizeof(DevCounters), hipHostMallocDefault));
    GR_HIP(hipMemsetAsync(d_bitmap, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_bits, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_ctr, 0, sizeof(DevCounters), stream));
    GR_HIP(hipMemcpyAsync(d_vd, vdata, sizeof(VD) * V, hipMemcpyHostToDevice, stream));
    GR_HIP(hipMemcpyAsync(d_vd_init, d_vd, sizeof(VD) * V, hipMemcpyDeviceToDevice, stream));
    if constexpr (!kEdgeEmpty) {
      GR_TRY(dalloc(&d_ed, E));
      if (E > 0) {
        const ED* src = static_cast<const ED*>(edata);
        ED* tmp = nullptr;
        if (!edata_on_device) {
          GR_TRY(dalloc(&tmp, E));
          GR_HIP(hipMemcpyAsync(tmp, edata, sizeof(ED) * E, hipMemcpyHostToDevice, stream));
          src = tmp;
        }
        k_permute_words<<<grid_1d(E, 256), 256, 0, stream>>>(
            reinterpret_cast<const uint32_t*>(src), g->csc_eid, E, (int)(sizeof(ED) / 4),
            reinterpret_cast<uint32_t*>(d_ed));
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(stream));
        if (tmp) (void)hipFree(tmp);
      }
    }
    GR_HIP(hipStreamSynchronize(stream));
    n_active = 0;
    w_edges = 0;
    iteration = 0;
    bound = true;
    return GR_OK;
  }

  const ED* edge_ptr() const {
    if constexpr (kEdgeEmpty)
      return reinterpret_cast<const ED*>(d_gtmp);  // never dereferenced
    else
      return d_ed;
  }

  int set_active(int32_t start, int32_t end) override {
    if (start < 0 || end < start || end > g->nV)
      return set_errorf(GR_ERR_INVALID, "setActive range [%d, %d) outside [0, %d)", start, end,
                        g->nV);
    const int n = end - start;
    GR_HIP(hipMemsetAsync(d_bitmap, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_bits, 0, sizeof(uint32_t) * (n_words + 1), stream));
    k_set_active_range<<<grid_1d((int64_t)n + 1, 256), 256, 0, stream>>>(start, n, g->off_in,
                                                                         d_active, d_escan, d_ctr);
    GR_HIP(hipGetLastError());
    n_active = n;
    w_edges = g->h_off_in[end] - g->h_off_in[start];
    iteration = 0;
    return GR_OK;
  }

  // Brackets the kernels of one phase with events on the engine's stream.
  void tick() {
    if (!timing) return;
    ev_open = nullptr;
    (void)hipEventCreate(&ev_open);
    (void)hipEventRecord(ev_open, stream);
  }
  void tock(int phase) {
    if (!timing || !ev_open) return;
    hipEvent_t e1 = nullptr;
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e1, stream);
    ev[phase].push_back(ev_open);
    ev[phase].push_back(e1);
    ev_open = nullptr;
  }

  int gather() override {
    if (!P::kHasGather || n_active == 0) return GR_OK;
    const int n = n_active, w = w_edges;
    const int n_tiles = (int)(((int64_t)n + w + kTile - 1) / kTile);
    k_merge_partition<<<grid_1d((int64_t)n_tiles + 1, 256), 256, 0, stream>>>(d_escan, n, w,
                                                                              n_tiles, d_part);
    tick();
    k_gather_tiles<P><<<n_tiles, kNT, 0, stream>>>(n, w, d_active, d_escan, d_part, g->off_in,
                                                    g->srcs, d_vd, edge_ptr(), d_gtmp, d_carry,
                                                    d_head);
    tock(kPhGather);
    k_gather_fixup<P><<<grid_1d(n_tiles, 256), 256, 0, stream>>>(n_tiles, d_carry, d_head, d_gtmp);
    GR_HIP(hipGetLastError());
    stats.active_vertices += n;
    stats.gathered_edges += w;
    return GR_OK;
  }

  int apply() override {
    if (n_active == 0) return GR_OK;
    const bool pull_possible = dir_mode != 1;
    tick();
    k_apply<P><<<grid_1d(n_active, 256), 256, 0, stream>>>(
        n_active, d_active, d_gtmp, P::kHasGather ? 1 : 0, d_vd, g->off_out, iteration, d_chg_v,
        d_chg_e, pull_possible ? d_chg_bits : nullptr, d_ctr);
    tock(kPhApply);
    GR_HIP(hipGetLastError());
    return GR_OK;
  }

  long long pull_threshold() const {
    return (long long)((double)g->nE / (dir_alpha > 0 ? dir_alpha : 1.0));
  }

  int scatter_activate(bool) override {
    // Program::scatter is a no-op for every compiled program (kHasScatter).
    static_assert(!P::kHasScatter, "non-trivial scatter is not compiled in");
    const long long thr = pull_threshold();
    tick();
    if (dir_mode != 1) {
      k_pull_activate<<<grid_1d(g->nV, 256), 256, 0, stream>>>(d_ctr, dir_mode, thr, g->nV,
                                                                g->off_in, g->srcs, d_chg_bits,
                                                                d_bitmap);
    }
    if (dir_mode != 2) {
      k_push_partition<<<grid_1d((int64_t)max_push_tiles + 1, 256), 256, 0, stream>>>(
          d_ctr, dir_mode, thr, d_chg_e, d_push_part, max_push_tiles);
      int grid = max_push_tiles < 8192 ? max_push_tiles : 8192;
      if (grid < 1) grid = 1;
      k_push_tiles<<<grid, kNT, 0, stream>>>(d_ctr, dir_mode, thr, d_push_part, d_chg_v, d_chg_e,
                                             g->off_out, g->dsts, d_bitmap);
    }
    tock(kPhScatter);
    GR_HIP(hipGetLastError());
    return GR_OK;
  }

  int next_iter() override {
    const int need_deg = P::kHasGather ? 1 : 0;
    tick();
    k_compact_count<<<n_cblocks, kCompactNT, 0, stream>>>(d_bitmap, n_words, g->off_in, need_deg,
                                                          d_blk);
    k_compact_scan<<<1, 1024, 0, stream>>>(d_blk, n_cblocks, d_ctr, need_deg ? d_escan : nullptr,
                                           dir_mode, pull_threshold());
    k_compact_write<<<n_cblocks, kCompactNT, 0, stream>>>(d_bitmap, n_words, g->off_in, need_deg,
                                                          d_blk, d_active, d_escan,
                                                          dir_mode != 1 ? d_chg_bits : nullptr);
    tock(kPhNext);
    GR_HIP(hipGetLastError());
    GR_HIP(hipMemcpyAsync(h_ctr, d_ctr, sizeof(DevCounters), hipMemcpyDeviceToHost, stream));
    GR_HIP(hipStreamSynchronize(stream));
    const unsigned long long c = h_ctr->chg_last;
    stats.changed_vertices += (int64_t)(c >> 32);
    stats.pushed_edges += (int64_t)(uint32_t)c;
    stats.pull_iterations += h_ctr->pull ? 1 : 0;
    if (n_active > 0) stats.iterations += 1;
    n_active = h_ctr->n_active;
    w_edges = h_ctr->w_edges;
    ++iteration;
    return GR_OK;
  }

  int get_results() override {
    GR_HIP(hipMemcpyAsync(host_vdata, d_vd, sizeof(VD) * g->nV, hipMemcpyDeviceToHost, stream));
    GR_HIP(hipStreamSynchronize(stream));
    return GR_OK;
  }

  int reset_vertex_data() override {
    GR_HIP(hipMemcpyAsync(d_vd, d_vd_init, sizeof(VD) * g->nV, hipMemcpyDeviceToDevice, stream));
    return GR_OK;
  }

  int read_vertex_data(void* out, size_t velem) override {
    if (velem != sizeof(VD)) return set_error(GR_ERR_INVALID, "vertex element size mismatch");
    GR_HIP(hipMemcpyAsync(out, d_vd, sizeof(VD) * g->nV, hipMemcpyDeviceToHost, stream));
    GR_HIP(hipStreamSynchronize(stream));
    return GR_OK;
  }

  static void sum_pairs(std::vector<hipEvent_t>& v, int64_t* launches, double* ms) {
    for (size_t i = 0; i + 1 < v.size(); i += 2) {
      float t = 0;
      if (hipEventElapsedTime(&t, v[i], v[i + 1]) == hipSuccess) *ms += t;
      *launches += 1;
    }
    for (hipEvent_t e : v) (void)hipEventDestroy(e);
    v.clear();
  }

  int sync_stats(gr_engine_stats* out) override {
    if (stream) GR_HIP(hipStreamSynchronize(stream));
    sum_pairs(ev[kPhGather], &stats.gather_launches, &stats.gather_ms);
    sum_pairs(ev[kPhScatter], &stats.scatter_launches, &stats.scatter_ms);
    sum_pairs(ev[kPhApply], &stats.apply_launches, &stats.apply_ms);
    sum_pairs(ev[kPhNext], &stats.next_launches, &stats.next_ms);
    if (out) *out = stats;
    return GR_OK;
  }
};

}  // namespace gr

struct gr_engine {
  int program = 0;
  gr::EngineBase* impl = nullptr;
};

using namespace gr;

namespace {

EngineBase* make_engine(int program) {
  switch (program) {
    case GR_PROGRAM_BFS: return new Engine<BFSProgram>();
    case GR_PROGRAM_SSSP: return new Engine<SSSPProgram>();
    case GR_PROGRAM_PAGERANK: return new Engine<PageRankProgram>();
    case GR_PROGRAM_CC: return new Engine<CCProgram>();
    default: return nullptr;
  }
}

int check(gr_engine* e, bool need_bound = true) {
  if (!e || !e->impl) return set_error(GR_ERR_INVALID, "engine is NULL");
  if (need_bound && !e->impl->bound) return set_error(GR_ERR_STATE, "engine has no graph (call setGraph)");
  return GR_OK;
}

}  // namespace

extern "C" int gr_engine_create(int program_id, gr_engine** out) {
  if (!out) return set_error(GR_ERR_INVALID, "out is NULL");
  EngineBase* impl = make_engine(program_id);
  if (!impl) return set_errorf(GR_ERR_UNSUPPORTED, "unknown program id %d", program_id);
  gr_engine* e = new gr_engine();
  e->program = program_id;
  e->impl = impl;
  *out = e;
  return GR_OK;
}

extern "C" int gr_engine_destroy(gr_engine* e) {
  if (!e) return GR_OK;
  delete e->impl;
  delete e;
  return GR_OK;
}

extern "C" int gr_engine_set_graph(gr_engine* e, int32_t n_vertices, void* vertex_data,
                                   size_t vertex_elem_size, int32_t n_edges,
                                   const void* edge_data, size_t edge_elem_size,
                                   const int32_t* srcs, const int32_t* dsts) {
  GR_TRY(check(e, false));
  if (vertex_elem_size != e->impl->vertex_size())
    return set_errorf(GR_ERR_INVALID, "vertex element size %zu != program VertexData size %zu",
                      vertex_elem_size, e->impl->vertex_size());
  if (e->impl->edge_size() && (!edge_data || edge_elem_size != e->impl->edge_size()))
    return set_errorf(GR_ERR_INVALID, "program needs edge data of %zu-byte elements",
                      e->impl->edge_size());
  if (!vertex_data) return set_error(GR_ERR_INVALID, "vertex data is NULL");
  gr_graph* g = nullptr;
  GR_TRY(gr_graph_from_edges(n_vertices, n_edges, srcs, dsts, &g));
  // bind() takes ownership of g once its argument checks (repeated above) pass
  return e->impl->bind(g, true, vertex_data, vertex_elem_size, edge_data, edge_elem_size, false);
}

extern "C" int gr_engine_bind_graph(gr_engine* e, gr_graph* g, void* vertex_data,
                                    size_t vertex_elem_size, const void* edge_data,
                                    size_t edge_elem_size, int edge_data_on_device) {
  GR_TRY(check(e, false));
  if (!g) return set_error(GR_ERR_INVALID, "graph is NULL");
  return e->impl->bind(g, false, vertex_data, vertex_elem_size, edge_data, edge_elem_size,
                       edge_data_on_device != 0);
}

extern "C" int gr_engine_set_active(gr_engine* e, int32_t start, int32_t end) {
  GR_TRY(check(e));
  return e->impl->set_active(start, end);
}

extern "C" int gr_engine_count_active(gr_engine* e, int32_t* n_active) {
  GR_TRY(check(e));
  if (!n_active) return set_error(GR_ERR_INVALID, "n_active is NULL");
  *n_active = e->impl->n_active;
  return GR_OK;
}

extern "C" int gr_engine_gather(gr_engine* e, int) {
  GR_TRY(check(e));
  return e->impl->gather();
}

extern "C" int gr_engine_apply(gr_engine* e) {
  GR_TRY(check(e));
  return e->impl->apply();
}

extern "C" int gr_engine_scatter_activate(gr_engine* e, int have_scatter) {
  GR_TRY(check(e));
  return e->impl->scatter_activate(have_scatter != 0);
}

extern "C" int gr_engine_next_iter(gr_engine* e, int32_t* n_active) {
  GR_TRY(check(e));
  GR_TRY(e->impl->next_iter());
  if (n_active) *n_active = e->impl->n_active;
  return GR_OK;
}

extern "C" int gr_engine_get_results(gr_engine* e) {
  GR_TRY(check(e));
  return e->impl->get_results();
}

extern "C" int gr_engine_run(gr_engine* e, int32_t* iterations) {
  GR_TRY(check(e));
  int32_t it = 0;
  while (e->impl->n_active) {
    GR_TRY(e->impl->gather());
    GR_TRY(e->impl->apply());
    GR_TRY(e->impl->scatter_activate(true));
    GR_TRY(e->impl->next_iter());
    ++it;
  }
  if (iterations) *iterations = it;
  return GR_OK;
}

extern "C" int gr_engine_set_iteration_count(gr_engine* e, int32_t iteration) {
  GR_TRY(check(e, false));
  e->impl->iteration = iteration;
  return GR_OK;
}

extern "C" int gr_engine_get_iteration_count(gr_engine* e, int32_t* iteration) {
  GR_TRY(check(e, false));
  if (!iteration) return set_error(GR_ERR_INVALID, "iteration is NULL");
  *iteration = e->impl->iteration;
  return GR_OK;
}

extern "C" int gr_engine_read_vertex_data(gr_engine* e, void* out, size_t vertex_elem_size) {
  GR_TRY(check(e));
  if (!out) return set_error(GR_ERR_INVALID, "out is NULL");
  return e->impl->read_vertex_data(out, vertex_elem_size);
}

extern "C" int gr_engine_reset_vertex_data(gr_engine* e) {
  GR_TRY(check(e));
  return e->impl->reset_vertex_data();
}

extern "C" int gr_engine_set_direction(gr_engine* e, int mode, double alpha) {
  GR_TRY(check(e, false));
  if (mode < 0 || mode > 2) return set_error(GR_ERR_INVALID, "direction mode must be 0, 1 or 2");
  e->impl->dir_mode = mode;
  if (alpha > 0) e->impl->dir_alpha = alpha;
  return GR_OK;
}

extern "C" int gr_engine_enable_timing(gr_engine* e, int on) {
  GR_TRY(check(e, false));
  e->impl->timing = on != 0;
  return GR_OK;
}

extern "C" int gr_engine_get_stats(gr_engine* e, gr_engine_stats* out) {
  GR_TRY(check(e, false));
  if (!out) return set_error(GR_ERR_INVALID, "out is NULL");
  return e->impl->sync_stats(out);
}

extern "C" int gr_engine_reset_stats(gr_engine* e) {
  GR_TRY(check(e, false));
  GR_TRY(e->impl->sync_stats(nullptr));
  e->impl->stats = gr_engine_stats{};
  return GR_OK;
}
