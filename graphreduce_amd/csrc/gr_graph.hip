// Device graph construction (CSR + CSC in HBM), device generators, and the
// runtime utilities of the C ABI.
//
// The reference builds CSR/CSC on the host with an unstable std::sort over the
// edge list (util.cuh:48-121, called from gpugas.h:1346-1373) and then streams
// shards to the GPU on every phase.  Here the whole graph is built once on the
// device with LSD radix sorts of packed (row, column) keys; the sort is stable,
// so parallel arcs keep their edge-list order and the layout is deterministic.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "gr_device.h"
#include "gr_hash.h"

namespace gr {

namespace {

constexpr int kBlock = 256;

inline int bits_for(int64_t n) {  // bits needed for values in [0, n)
  int b = 1;
  while (b < 62 && ((int64_t)1 << b) < n) ++b;
  return b;
}

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (1 << 30)) g = 1 << 30;
  return (unsigned)g;
}

__global__ void k_pack_keys(const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                            int64_t n, int shift, uint64_t* __restrict__ keys,
                            int32_t* __restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = ((uint64_t)(uint32_t)rows[i] << shift) | (uint32_t)cols[i];
  vals[i] = (int32_t)i;
}

__global__ void k_iota(int32_t* __restrict__ v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}

__global__ void k_swap_keys(const uint64_t* __restrict__ in, int64_t n, int shift,
                            uint64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t mask = (1ULL << shift) - 1;
  uint64_t k = in[i];
  out[i] = ((k & mask) << shift) | (k >> shift);
}

// Column part of sorted keys, and the row offsets: offsets[v] = first position
// whose row is >= v.  Positions fill the gaps before their own row; the tail
// (rows after the last key) is filled by k_offsets_tail.
__global__ void k_split_sorted(const uint64_t* __restrict__ keys, int64_t n, int shift,
                               int32_t* __restrict__ cols, int32_t* __restrict__ offsets) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t mask = (1ULL << shift) - 1;
  uint64_t k = keys[i];
  int32_t r = (int32_t)(k >> shift);
  cols[i] = (int32_t)(k & mask);
  int32_t pr = (i == 0) ? -1 : (int32_t)(keys[i - 1] >> shift);
  for (int32_t v = pr + 1; v <= r; ++v) offsets[v] = (int32_t)i;
}

__global__ void k_offsets_tail(const uint64_t* __restrict__ keys, int64_t n, int shift,
                               int32_t nV, int32_t* __restrict__ offsets) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nV) return;
  int32_t last = (n == 0) ? -1 : (int32_t)(keys[n - 1] >> shift);
  if (v > last) offsets[v] = (int32_t)n;
}

__global__ void k_rmat_keys(uint64_t base, int64_t m, int scale, RmatThresholds t, int sym,
                            uint64_t* __restrict__ keys) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  uint32_t s, d;
  rmat_edge(base, (uint64_t)k, scale, t, &s, &d);
  keys[k] = ((uint64_t)s << scale) | d;
  if (sym) keys[m + k] = ((uint64_t)d << scale) | s;
}

__global__ void k_mark_unique(const uint64_t* __restrict__ keys, int64_t n, int shift,
                              uint8_t* __restrict__ keep) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t mask = (1ULL << shift) - 1;
  uint64_t k = keys[i];
  bool self = (k >> shift) == (k & mask);
  bool dup = (i > 0) && keys[i - 1] == k;
  keep[i] = (!self && !dup) ? 1 : 0;
}

__global__ void k_hash_weights(uint64_t base, int64_t n, uint32_t lo, int span_log2,
                               int32_t* __restrict__ w) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = (int32_t)hash_weight(base, (uint64_t)i, lo, span_log2);
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(int64_t n) {
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
      p = nullptr;
      return set_errorf(GR_ERR_NOMEM, "hipMalloc(%lld bytes) failed: %s",
                        (long long)(sizeof(T) * n), hipGetErrorString(e));
    }
    return GR_OK;
  }
  T* release() {
    T* r = p;
    p = nullptr;
    return r;
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};

// Sorts (keys, vals) by the low `bits` bits of the keys; results in *_out.
int radix_sort_pairs(uint64_t* keys_in, uint64_t* keys_out, int32_t* vals_in,
                     int32_t* vals_out, int64_t n, int bits) {
  if (n == 0) return GR_OK;
  size_t temp = 0;
  GR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys_in, keys_out, vals_in,
                                            vals_out, (int)n, 0, bits));
  DevBuf<uint8_t> t;
  GR_TRY(t.alloc((int64_t)temp));
  GR_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, temp, keys_in, keys_out, vals_in, vals_out,
                                            (int)n, 0, bits));
  return GR_OK;
}

// One side (CSR or CSC) from keys row << shift | col (unsorted) with values.
// On return: *off [V+1], *adj [E], *eid [E] are fresh device arrays.
int build_side(int32_t nV, int64_t n, int shift, uint64_t* keys, int32_t* vals,
               int32_t** off, int32_t** adj, int32_t** eid) {
  DevBuf<uint64_t> ksorted;
  DevBuf<int32_t> vsorted, o, a;
  GR_TRY(ksorted.alloc(n));
  GR_TRY(vsorted.alloc(n));
  GR_TRY(o.alloc((int64_t)nV + 1));
  GR_TRY(a.alloc(n));
  GR_TRY(radix_sort_pairs(keys, ksorted.p, vals, vsorted.p, n, shift + bits_for(nV)));
  if (n > 0) {
    k_split_sorted<<<grid_for(n), kBlock>>>(ksorted.p, n, shift, a.p, o.p);
    GR_HIP(hipGetLastError());
  }
  k_offsets_tail<<<grid_for((int64_t)nV + 1), kBlock>>>(ksorted.p, n, shift, nV, o.p);
  GR_HIP(hipGetLastError());
  *off = o.release();
  *adj = a.release();
  *eid = vsorted.release();
  return GR_OK;
}

int finish_graph(gr_graph* g) {
  g->h_off_out.resize((size_t)g->nV + 1);
  g->h_off_in.resize((size_t)g->nV + 1);
  GR_HIP(hipMemcpy(g->h_off_out.data(), g->off_out, sizeof(int32_t) * (g->nV + 1),
                   hipMemcpyDeviceToHost));
  GR_HIP(hipMemcpy(g->h_off_in.data(), g->off_in, sizeof(int32_t) * (g->nV + 1),
                   hipMemcpyDeviceToHost));
  return GR_OK;
}

}  // namespace

int graph_free(gr_graph* g) {
  if (!g) return GR_OK;
  if (g->ordered) graph_free(g->ordered);
  for (gr_blocked* b : {g->blocked, g->blocked_half}) {
    if (!b) continue;
    for (void* p : {(void*)b->brow, (void*)b->bnr, (void*)b->boff, (void*)b->bpk,
                    (void*)b->gbase, (void*)b->bpos})
      if (p) (void)hipFree(p);
    delete b;
  }
  for (int32_t* p : {g->off_out, g->dsts, g->off_in, g->srcs, g->csr_eid, g->csc_eid, g->perm,
                     g->inv})
    if (p) (void)hipFree(p);
  delete g;
  return GR_OK;
}

namespace {

__global__ void k_degree_keys(int32_t nV, const int32_t* __restrict__ off_out,
                              const int32_t* __restrict__ off_in, uint64_t* __restrict__ keys,
                              int32_t* __restrict__ ids) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const uint32_t deg = (uint32_t)(off_out[v + 1] - off_out[v]) + (uint32_t)(off_in[v + 1] - off_in[v]);
  keys[v] = (uint64_t)(0xFFFFFFFFu - deg);  // ascending key = descending degree
  ids[v] = (int32_t)v;
}

__global__ void k_inverse(int32_t nV, const int32_t* __restrict__ perm, int32_t* __restrict__ inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nV) inv[perm[i]] = (int32_t)i;
}

// rowid[off[v]] = v for every non-empty row (rowid zeroed before)
__global__ void k_row_marks(int32_t nV, const int32_t* __restrict__ off, int32_t* __restrict__ rowid) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nV && off[v + 1] > off[v]) rowid[off[v]] = (int32_t)v;
}

// The edge list in the new ids, by edge-list position: CSR position p is
// edge eid[p] (identity when eid is nullptr) from row rowid[p] to adj[p].
__global__ void k_relabel_edges(int64_t nE, const int32_t* __restrict__ rowid,
                                const int32_t* __restrict__ adj, const int32_t* __restrict__ eid,
                                const int32_t* __restrict__ inv, int32_t* __restrict__ s_new,
                                int32_t* __restrict__ d_new) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nE) return;
  const int64_t i = eid ? (int64_t)eid[p] : p;
  s_new[i] = inv[rowid[p]];
  d_new[i] = inv[adj[p]];
}

// Row of every CSR position (max-scan over the row starts).
int row_ids(int32_t nV, int64_t nE, const int32_t* off, int32_t* rowid) {
  if (nE == 0) return GR_OK;
  DevBuf<int32_t> marks;
  GR_TRY(marks.alloc(nE));
  GR_HIP(hipMemset(marks.p, 0, sizeof(int32_t) * (size_t)nE));
  k_row_marks<<<grid_for(nV), kBlock>>>(nV, off, marks.p);
  GR_HIP(hipGetLastError());
  size_t temp = 0;
  GR_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, temp, marks.p, rowid, hipcub::Max(), (int)nE));
  DevBuf<uint8_t> t;
  GR_TRY(t.alloc((int64_t)temp));
  GR_HIP(hipcub::DeviceScan::InclusiveScan(t.p, temp, marks.p, rowid, hipcub::Max(), (int)nE));
  return GR_OK;
}

std::mutex g_order_mu;

}  // namespace

int graph_from_device_edges(int32_t nV, int64_t nE, const int32_t* d_src,
                            const int32_t* d_dst, gr_graph** out);

int graph_ordered(gr_graph* g, gr_graph** out) {
  std::lock_guard<std::mutex> lk(g_order_mu);
  if (g->ordered) {
    *out = g->ordered;
    return GR_OK;
  }
  const int32_t nV = g->nV;
  const int64_t nE = g->nE;
  DevBuf<int32_t> perm, inv;
  GR_TRY(perm.alloc(nV));
  GR_TRY(inv.alloc(nV));
  if (nV > 0) {
    DevBuf<uint64_t> keys, ksorted;
    DevBuf<int32_t> ids;
    GR_TRY(keys.alloc(nV));
    GR_TRY(ksorted.alloc(nV));
    GR_TRY(ids.alloc(nV));
    k_degree_keys<<<grid_for(nV), kBlock>>>(nV, g->off_out, g->off_in, keys.p, ids.p);
    GR_HIP(hipGetLastError());
    GR_TRY(radix_sort_pairs(keys.p, ksorted.p, ids.p, perm.p, nV, 32));  // stable
    k_inverse<<<grid_for(nV), kBlock>>>(nV, perm.p, inv.p);
    GR_HIP(hipGetLastError());
  }
  // Rebuilt from the renumbered edge list (same edge-list positions, so the
  // eid maps still index the caller's edge data); rows come out sorted by the
  // new neighbour ids, which puts the hubs of every row next to each other.
  gr_graph* o = nullptr;
  {
    DevBuf<int32_t> rowid, s_new, d_new;
    GR_TRY(rowid.alloc(nE));
    GR_TRY(s_new.alloc(nE));
    GR_TRY(d_new.alloc(nE));
    GR_TRY(row_ids(nV, nE, g->off_out, rowid.p));
    if (nE > 0) {
      k_relabel_edges<<<grid_for(nE), kBlock>>>(nE, rowid.p, g->dsts, g->csr_eid, inv.p, s_new.p,
                                                d_new.p);
      GR_HIP(hipGetLastError());
    }
    GR_TRY(graph_from_device_edges(nV, nE, s_new.p, d_new.p, &o));
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    graph_free(o);
    return set_error(GR_ERR_HIP, "ordered graph build failed");
  }
  g->ordered = o;
  g->perm = perm.release();
  g->inv = inv.release();
  *out = o;
  return GR_OK;
}

namespace {

// key[j] = block of CSC position j's row << shift | its source; val[j] = j.
__global__ void k_block_keys(int64_t nE, const int32_t* __restrict__ rowid,
                             const int32_t* __restrict__ row_block,
                             const int32_t* __restrict__ srcs, int shift,
                             uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nE) return;
  keys[j] = ((uint64_t)(uint32_t)row_block[rowid[j]] << shift) | (uint32_t)srcs[j];
  vals[j] = (int32_t)j;
}

// Sorted keys -> source, row within its block, CSC position.
__global__ void k_block_split(int64_t nE, const uint64_t* __restrict__ keys,
                              const int32_t* __restrict__ vals, int shift,
                              const int32_t* __restrict__ rowid, const int32_t* __restrict__ brow,
                              int32_t* __restrict__ bsrc, uint16_t* __restrict__ brl,
                              int32_t* __restrict__ bpos) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nE) return;
  const uint64_t key = keys[k];
  const int32_t j = vals[k];
  bsrc[k] = (int32_t)(key & ((1ull << shift) - 1));
  brl[k] = (uint16_t)(rowid[j] - brow[(int32_t)(key >> shift)]);
  bpos[k] = j;
}

// Slots of sorted in-edge k in the packed layout: its own, and one pad slot
// per kPackStep of source distance beyond the first from the block's previous
// in-edge.  size[nE] = 0 closes the array for the exclusive scan.
__global__ void k_pack_sizes(int64_t nE, const uint64_t* __restrict__ keys, int shift,
                             const int32_t* __restrict__ bsrc, int32_t* __restrict__ size) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nE) return;
  if (k == nE) {
    size[k] = 0;
    return;
  }
  int32_t n = 1;
  if (k > 0 && (keys[k] >> shift) == (keys[k - 1] >> shift)) {
    const int32_t gap = bsrc[k] - bsrc[k - 1];  // >= 0: sorted by source within the block
    if (gap > kPackStep) n += (gap - 1) / kPackStep;
  }
  size[k] = n;
}

// out[i] = in[idx[i]].
__global__ void k_take(int n, const int32_t* __restrict__ idx, const int32_t* __restrict__ in,
                       int32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

// Every slot a pad of step 0, no in-edge.
__global__ void k_pack_fill(int64_t n, uint32_t* __restrict__ bpk, int32_t* __restrict__ bpos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bpk[i] = kPackPad;
  bpos[i] = -1;
}

__device__ __forceinline__ void pack_slot(int64_t s, int32_t step, uint32_t row, int32_t src,
                                          int32_t pos, uint32_t* __restrict__ bpk,
                                          int32_t* __restrict__ gbase,
                                          int32_t* __restrict__ bpos) {
  if (s % kPackGroup == 0) {  // a group's first slot: its source is the base
    gbase[s / kPackGroup] = src;
    step = 0;
  }
  bpk[s] = ((uint32_t)step << 16) | row;
  bpos[s] = pos;
}

// Sorted in-edge k and the pad slots before it, into the packed layout:
// in-edge k of block b starts its run at slot pstart[b] + scan[k] - scan of
// the block's first in-edge (estart[b]).
__global__ void k_pack_scatter(int64_t nE, const uint64_t* __restrict__ keys, int shift,
                               const int32_t* __restrict__ bsrc, const uint16_t* __restrict__ brl,
                               const int32_t* __restrict__ pos, const int32_t* __restrict__ scan,
                               const int32_t* __restrict__ estart,
                               const int32_t* __restrict__ pstart, uint32_t* __restrict__ bpk,
                               int32_t* __restrict__ gbase, int32_t* __restrict__ bpos) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nE) return;
  const int32_t b = (int32_t)(keys[k] >> shift);
  const bool first = k == 0 || (keys[k - 1] >> shift) != (uint64_t)b;
  const int64_t s0 = (int64_t)pstart[b] + (scan[k] - estart[b]);
  const int32_t pads = scan[k + 1] - scan[k] - 1;
  const int32_t prev = first ? bsrc[k] : bsrc[k - 1];
  for (int32_t i = 0; i < pads; ++i)
    pack_slot(s0 + i, kPackStep, kPackPad, prev + (i + 1) * kPackStep, -1, bpk, gbase, bpos);
  const int32_t step = first ? 0 : bsrc[k] - prev - pads * kPackStep;
  pack_slot(s0 + pads, step, brl[k], bsrc[k], pos[k], bpk, gbase, bpos);
}

// Decodes every group of the packed layout as the blocked gather does and
// compares each slot with the CSC: err[0] counts mismatches (a step in a
// group's first slot, a pad slot naming an in-edge, a source or row other than
// the CSC's), err[1] the slots that name an in-edge.
__global__ void k_pack_check(int64_t n_groups, const uint32_t* __restrict__ bpk,
                             const int32_t* __restrict__ gbase, const int32_t* __restrict__ bpos,
                             const int32_t* __restrict__ pstart, int nb,
                             const int32_t* __restrict__ brow, const int32_t* __restrict__ srcs,
                             const int32_t* __restrict__ rowid, int64_t nE,
                             unsigned long long* __restrict__ err) {
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n_groups) return;
  const int64_t s0 = gi * kPackGroup;
  int lo = 0, hi = nb;  // the block: last pstart[b] <= s0
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if ((int64_t)pstart[mid] <= s0) lo = mid; else hi = mid;
  }
  unsigned long long bad = 0, real = 0;
  int32_t src = gbase[gi];
  for (int i = 0; i < kPackGroup; ++i) {
    const uint32_t w = bpk[s0 + i];
    if (i == 0) bad += (w >> 16) != 0u;
    else src += (int32_t)(w >> 16);
    const uint32_t row = w & 0xFFFFu;
    const int32_t j = bpos[s0 + i];
    if (row == kPackPad) {
      bad += j != -1;
      continue;
    }
    ++real;
    if (j < 0 || (int64_t)j >= nE || srcs[j] != src || rowid[j] - brow[lo] != (int32_t)row) ++bad;
  }
  if (bad) atomicAdd(&err[0], bad);
  if (real) atomicAdd(&err[1], real);
}

std::mutex g_block_mu;

}  // namespace

int graph_blocked(gr_graph* g, int max_rows, const gr_blocked** out) {
  std::lock_guard<std::mutex> lk(g_block_mu);
  if (max_rows != kBlockRows && max_rows != kBlockRowsWide)
    return set_errorf(GR_ERR_INVALID, "graph_blocked: %d rows per block", max_rows);
  gr_blocked*& slot = max_rows == kBlockRows ? g->blocked : g->blocked_half;
  if (slot) {
    *out = slot;
    return GR_OK;
  }
  const int32_t nV = g->nV;
  const int64_t nE = g->nE;
  const std::vector<int32_t>& off = g->h_off_in;
  // blocks of consecutive rows: at most max_rows rows, at most edge_target
  // in-edges unless a single row has more
  int32_t last = nV;
  while (last > 0 && off[last] == off[last - 1]) --last;
  // about two blocks per compute unit: the blocks of one launch then share the
  // sweep through the source ids (measured at RMAT-24: 686 blocks of <= 2^19
  // in-edges 0.74 ms, 1376 of <= 2^18 0.91 ms, 459 of <= 2^20 0.84 ms)
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  const int64_t edge_target =
      std::max<int64_t>(4096, std::min<int64_t>(kBlockEdges, nE / (2 * (int64_t)cus)));
  std::vector<int32_t> brow, bnr, boff, row_block((size_t)std::max(nV, 1), 0);
  for (int32_t r = 0; r < last;) {
    const int64_t target = (int64_t)off[r] + edge_target;
    int32_t r1 = (int32_t)(std::upper_bound(off.begin() + r + 1, off.begin() + last + 1,
                                            (int32_t)std::min<int64_t>(target, INT32_MAX)) -
                           off.begin()) - 1;
    r1 = std::min(std::min(r1, r + max_rows), last);
    if (r1 <= r) r1 = r + 1;
    for (int32_t v = r; v < r1; ++v) row_block[v] = (int32_t)brow.size();
    brow.push_back(r);
    bnr.push_back(r1 - r);
    boff.push_back(off[r]);
    r = r1;
  }
  boff.push_back(off[last]);
  const int32_t nb = (int32_t)brow.size();
  // Slot bound: every in-edge, the pads of a block (its source range over
  // kPackStep, at most nV / kPackStep + 1) and a group of fill per block.
  const int64_t slot_bound =
      nE + (int64_t)nb * ((int64_t)nV / kPackStep + 1 + kPackGroup) + kPackGroup;
  if (slot_bound >= (int64_t)INT32_MAX)
    return set_errorf(GR_ERR_UNSUPPORTED, "graph_blocked: %lld slots exceed int32",
                      (long long)slot_bound);
  gr_blocked* b = new gr_blocked();
  b->n_blocks = nb;
  b->rows_end = last;
  DevBuf<int32_t> d_brow, d_bnr, d_boff, d_rowblk, d_gbase, d_bpos;
  DevBuf<uint32_t> d_bpk;
  int rc = GR_OK;
  auto fail = [&](int code) {
    delete b;
    return code;
  };
  if ((rc = d_brow.alloc(nb)) || (rc = d_bnr.alloc(nb)) || (rc = d_boff.alloc(nb + 1)) ||
      (rc = d_rowblk.alloc(nV)))
    return fail(rc);
  if (nb > 0) {
    if (hipMemcpy(d_brow.p, brow.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice) ||
        hipMemcpy(d_bnr.p, bnr.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice))
      return fail(set_error(GR_ERR_HIP, "graph_blocked: block table upload"));
  }
  if (nV > 0 && hipMemcpy(d_rowblk.p, row_block.data(), sizeof(int32_t) * nV,
                          hipMemcpyHostToDevice))
    return fail(set_error(GR_ERR_HIP, "graph_blocked: block table upload"));
  std::vector<int32_t> pstart((size_t)nb + 1, 0);
  if (nE > 0) {
    const int shift = bits_for(nV);
    DevBuf<uint64_t> keys, ksorted;
    DevBuf<int32_t> vals, vsorted, rowid, bsrc, size, scan, eidx, estart, d_pstart;
    DevBuf<uint16_t> brl;
    if ((rc = keys.alloc(nE)) || (rc = ksorted.alloc(nE)) || (rc = vals.alloc(nE)) ||
        (rc = vsorted.alloc(nE)) || (rc = rowid.alloc(nE)))
      return fail(rc);
    if ((rc = row_ids(nV, nE, g->off_in, rowid.p))) return fail(rc);
    k_block_keys<<<grid_for(nE), kBlock>>>(nE, rowid.p, d_rowblk.p, g->srcs, shift, keys.p,
                                           vals.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error(GR_ERR_HIP, "k_block_keys"));
    if ((rc = radix_sort_pairs(keys.p, ksorted.p, vals.p, vsorted.p, nE, shift + bits_for(nb))))
      return fail(rc);
    keys.reset();
    vals.reset();
    // sorted in-edges: source, row in block, CSC position
    DevBuf<int32_t> pos;
    if ((rc = bsrc.alloc(nE)) || (rc = brl.alloc(nE)) || (rc = pos.alloc(nE))) return fail(rc);
    k_block_split<<<grid_for(nE), kBlock>>>(nE, ksorted.p, vsorted.p, shift, rowid.p, d_brow.p,
                                            bsrc.p, brl.p, pos.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error(GR_ERR_HIP, "k_block_split"));
    vsorted.reset();
    // slots per in-edge, their exclusive scan (nE + 1 entries), the scan at
    // each block's first in-edge
    if ((rc = size.alloc(nE + 1)) || (rc = scan.alloc(nE + 1)) || (rc = eidx.alloc(nb + 1)) ||
        (rc = estart.alloc(nb + 1)))
      return fail(rc);
    k_pack_sizes<<<grid_for(nE + 1), kBlock>>>(nE, ksorted.p, shift, bsrc.p, size.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error(GR_ERR_HIP, "k_pack_sizes"));
    {
      size_t temp = 0;
      if (hipcub::DeviceScan::ExclusiveSum(nullptr, temp, size.p, scan.p, (int)(nE + 1)))
        return fail(set_error(GR_ERR_HIP, "graph_blocked: slot scan"));
      DevBuf<uint8_t> t;
      if ((rc = t.alloc((int64_t)temp))) return fail(rc);
      if (hipcub::DeviceScan::ExclusiveSum(t.p, temp, size.p, scan.p, (int)(nE + 1)))
        return fail(set_error(GR_ERR_HIP, "graph_blocked: slot scan"));
    }
    if (hipMemcpy(eidx.p, boff.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice))
      return fail(set_error(GR_ERR_HIP, "graph_blocked: block table upload"));
    k_take<<<grid_for(nb + 1), kBlock>>>(nb + 1, eidx.p, scan.p, estart.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error(GR_ERR_HIP, "k_take"));
    std::vector<int32_t> h_estart((size_t)nb + 1);
    if (hipMemcpy(h_estart.data(), estart.p, sizeof(int32_t) * (nb + 1), hipMemcpyDeviceToHost))
      return fail(set_error(GR_ERR_HIP, "graph_blocked: slot counts"));
    // every block starts on a group; its last group is filled up
    for (int32_t k = 0; k < nb; ++k) {
      const int64_t len = (int64_t)h_estart[k + 1] - h_estart[k];
      pstart[k + 1] = (int32_t)((pstart[k] + len + kPackGroup - 1) / kPackGroup * kPackGroup);
    }
    b->n_slots = pstart[nb];
    if ((rc = d_pstart.alloc(nb + 1)) || (rc = d_bpk.alloc(b->n_slots)) ||
        (rc = d_bpos.alloc(b->n_slots)) || (rc = d_gbase.alloc(b->n_slots / kPackGroup)))
      return fail(rc);
    if (hipMemcpy(d_pstart.p, pstart.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice))
      return fail(set_error(GR_ERR_HIP, "graph_blocked: block table upload"));
    k_pack_fill<<<grid_for(b->n_slots), kBlock>>>(b->n_slots, d_bpk.p, d_bpos.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error(GR_ERR_HIP, "k_pack_fill"));
    k_pack_scatter<<<grid_for(nE), kBlock>>>(nE, ksorted.p, shift, bsrc.p, brl.p, pos.p, scan.p,
                                             estart.p, d_pstart.p, d_bpk.p, d_gbase.p, d_bpos.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error(GR_ERR_HIP, "k_pack_scatter"));
    DevBuf<unsigned long long> err;
    if ((rc = err.alloc(2))) return fail(rc);
    if (hipMemset(err.p, 0, 2 * sizeof(unsigned long long)))
      return fail(set_error(GR_ERR_HIP, "graph_blocked: check counters"));
    const int64_t n_groups = b->n_slots / kPackGroup;
    k_pack_check<<<grid_for(n_groups), kBlock>>>(n_groups, d_bpk.p, d_gbase.p, d_bpos.p,
                                                 d_pstart.p, nb, d_brow.p, g->srcs, rowid.p, nE,
                                                 err.p);
    if (hipGetLastError() != hipSuccess) return fail(set_error(GR_ERR_HIP, "k_pack_check"));
    unsigned long long h_err[2] = {0, 0};
    if (hipMemcpy(h_err, err.p, sizeof(h_err), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(set_error(GR_ERR_HIP, "blocked CSC build failed"));
    if (h_err[0] != 0 || h_err[1] != (unsigned long long)nE)
      return fail(set_errorf(GR_ERR_HIP,
                             "graph_blocked: packed layout check: %llu bad slots, %llu of %lld "
                             "in-edges placed",
                             h_err[0], h_err[1], (long long)nE));
  }
  if (hipMemcpy(d_boff.p, pstart.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice))
    return fail(set_error(GR_ERR_HIP, "graph_blocked: block table upload"));
  b->brow = d_brow.release();
  b->bnr = d_bnr.release();
  b->boff = d_boff.release();
  b->bpk = d_bpk.release();
  b->gbase = d_gbase.release();
  b->bpos = d_bpos.release();
  slot = b;
  *out = b;
  return GR_OK;
}

int graph_from_device_edges(int32_t nV, int64_t nE, const int32_t* d_src,
                            const int32_t* d_dst, gr_graph** out) {
  const int shift = bits_for(nV);
  gr_graph* g = new gr_graph();
  g->nV = nV;
  g->nE = nE;
  int rc = GR_OK;
  {
    DevBuf<uint64_t> keys;
    DevBuf<int32_t> vals;
    if ((rc = keys.alloc(nE)) || (rc = vals.alloc(nE))) goto fail;
    if (nE > 0) {
      k_pack_keys<<<grid_for(nE), kBlock>>>(d_src, d_dst, nE, shift, keys.p, vals.p);
      if (hipGetLastError() != hipSuccess) { rc = set_error(GR_ERR_HIP, "k_pack_keys"); goto fail; }
    }
    if ((rc = build_side(nV, nE, shift, keys.p, vals.p, &g->off_out, &g->dsts, &g->csr_eid)))
      goto fail;
    if (nE > 0) {
      k_pack_keys<<<grid_for(nE), kBlock>>>(d_dst, d_src, nE, shift, keys.p, vals.p);
      if (hipGetLastError() != hipSuccess) { rc = set_error(GR_ERR_HIP, "k_pack_keys"); goto fail; }
    }
    if ((rc = build_side(nV, nE, shift, keys.p, vals.p, &g->off_in, &g->srcs, &g->csc_eid)))
      goto fail;
  }
  if ((rc = finish_graph(g))) goto fail;
  *out = g;
  return GR_OK;
fail:
  graph_free(g);
  return rc;
}

// Graph from sorted, de-duplicated keys src << shift | dst (edge-list order =
// CSR order, so csr_eid stays the identity).
int graph_from_sorted_keys(int32_t nV, int64_t nE, int shift, uint64_t* d_keys,
                           gr_graph** out) {
  gr_graph* g = new gr_graph();
  g->nV = nV;
  g->nE = nE;
  int rc = GR_OK;
  {
    DevBuf<int32_t> off, adj;
    if ((rc = off.alloc((int64_t)nV + 1)) || (rc = adj.alloc(nE))) goto fail;
    if (nE > 0) {
      k_split_sorted<<<grid_for(nE), kBlock>>>(d_keys, nE, shift, adj.p, off.p);
    }
    k_offsets_tail<<<grid_for((int64_t)nV + 1), kBlock>>>(d_keys, nE, shift, nV, off.p);
    if (hipGetLastError() != hipSuccess) { rc = set_error(GR_ERR_HIP, "csr split"); goto fail; }
    g->off_out = off.release();
    g->dsts = adj.release();
    DevBuf<uint64_t> swapped;
    DevBuf<int32_t> vals;
    if ((rc = swapped.alloc(nE)) || (rc = vals.alloc(nE))) goto fail;
    if (nE > 0) {
      k_swap_keys<<<grid_for(nE), kBlock>>>(d_keys, nE, shift, swapped.p);
      k_iota<<<grid_for(nE), kBlock>>>(vals.p, nE);
      if (hipGetLastError() != hipSuccess) { rc = set_error(GR_ERR_HIP, "csc keys"); goto fail; }
    }
    if ((rc = build_side(nV, nE, shift, swapped.p, vals.p, &g->off_in, &g->srcs, &g->csc_eid)))
      goto fail;
  }
  if ((rc = finish_graph(g))) goto fail;
  *out = g;
  return GR_OK;
fail:
  graph_free(g);
  return rc;
}

}  // namespace gr

using namespace gr;

// ------------------------------------------------------------------ runtime

extern "C" int gr_device_count(int* count) {
  if (!count) return set_error(GR_ERR_INVALID, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return GR_OK;
}

extern "C" int gr_device_info(char* name, size_t name_len, int* cu_count, size_t* total_bytes) {
  hipDeviceProp_t p;
  GR_HIP(hipGetDeviceProperties(&p, 0));
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (total_bytes) *total_bytes = p.totalGlobalMem;
  return GR_OK;
}

extern "C" int gr_device_synchronize(void) {
  GR_HIP(hipDeviceSynchronize());
  return GR_OK;
}

extern "C" int gr_device_alloc(size_t bytes, void** device_ptr) {
  if (!device_ptr) return set_error(GR_ERR_INVALID, "device_ptr is NULL");
  hipError_t e = hipMalloc(device_ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return set_errorf(GR_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return GR_OK;
}

extern "C" int gr_device_free(void* device_ptr) {
  if (device_ptr) GR_HIP(hipFree(device_ptr));
  return GR_OK;
}

namespace {
__global__ void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" int gr_device_copy(void* dst, const void* src, size_t bytes, float* ms) {
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return set_error(GR_ERR_INVALID, "gr_device_copy: 16-byte aligned sizes and pointers only");
  const int64_t n = (int64_t)(bytes / 16);
  hipEvent_t e0, e1;
  GR_HIP(hipEventCreate(&e0));
  GR_HIP(hipEventCreate(&e1));
  GR_HIP(hipEventRecord(e0, 0));
  k_stream_copy<<<4096, 256>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
  GR_HIP(hipGetLastError());
  GR_HIP(hipEventRecord(e1, 0));
  GR_HIP(hipEventSynchronize(e1));
  float t = 0;
  GR_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return GR_OK;
}

namespace {
// One read of dist[srcs[e]] per CSC slot (source ids streamed non-temporally,
// as the gather loads them); each thread keeps a minimum so the reads count.
__global__ __launch_bounds__(256) void k_source_reads(const int32_t* __restrict__ srcs,
                                                      int64_t n, const int32_t* __restrict__ dist,
                                                      int32_t* __restrict__ out) {
  constexpr int64_t kPer = 4096;
  const int64_t b0 = (int64_t)blockIdx.x * kPer;
  const int64_t b1 = b0 + kPer < n ? b0 + kPer : n;
  int32_t acc = INT_MAX;
  for (int64_t e = b0 + threadIdx.x; e < b1; e += blockDim.x)
    acc = min(acc, dist[__builtin_nontemporal_load(srcs + e)]);
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}
}  // namespace

extern "C" int gr_graph_source_read_probe(gr_graph* g, int ordered, int repeats, float* ms) {
  if (!g || repeats < 1) return set_error(GR_ERR_INVALID, "gr_graph_source_read_probe: bad arguments");
  gr_graph* lay = g;
  if (ordered) GR_TRY(gr::graph_ordered(g, &lay));
  const int64_t E = lay->nE;
  const unsigned blocks = (unsigned)((E + 4095) / 4096 > 0 ? (E + 4095) / 4096 : 1);
  DevBuf<int32_t> dist, out;
  GR_TRY(dist.alloc(lay->nV));
  GR_TRY(out.alloc((int64_t)blocks * 256));
  GR_HIP(hipMemset(dist.p, 0, sizeof(int32_t) * (size_t)(lay->nV > 0 ? lay->nV : 1)));
  k_source_reads<<<blocks, 256>>>(lay->srcs, E, dist.p, out.p);  // warm
  GR_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  GR_HIP(hipEventCreate(&e0));
  GR_HIP(hipEventCreate(&e1));
  GR_HIP(hipEventRecord(e0, 0));
  for (int r = 0; r < repeats; ++r) k_source_reads<<<blocks, 256>>>(lay->srcs, E, dist.p, out.p);
  GR_HIP(hipGetLastError());
  GR_HIP(hipEventRecord(e1, 0));
  GR_HIP(hipEventSynchronize(e1));
  float t = 0;
  GR_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t / repeats;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return GR_OK;
}

extern "C" int gr_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  if (bytes) GR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GR_OK;
}

extern "C" int gr_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  if (bytes) GR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GR_OK;
}

// ------------------------------------------------------------------- graph

extern "C" int gr_graph_from_edges(int32_t n_vertices, int32_t n_edges, const int32_t* srcs,
                                   const int32_t* dsts, gr_graph** out) {
  if (!out || n_vertices < 1 || n_edges < 0 || (n_edges > 0 && (!srcs || !dsts)))
    return set_error(GR_ERR_INVALID, "gr_graph_from_edges: bad arguments");
  for (int32_t i = 0; i < n_edges; ++i) {
    if ((uint32_t)srcs[i] >= (uint32_t)n_vertices || (uint32_t)dsts[i] >= (uint32_t)n_vertices)
      return set_errorf(GR_ERR_INVALID, "gr_graph_from_edges: edge %d (%d -> %d) out of range",
                        i, srcs[i], dsts[i]);
  }
  DevBuf<int32_t> s, d;
  GR_TRY(s.alloc(n_edges));
  GR_TRY(d.alloc(n_edges));
  if (n_edges) {
    GR_HIP(hipMemcpy(s.p, srcs, sizeof(int32_t) * n_edges, hipMemcpyHostToDevice));
    GR_HIP(hipMemcpy(d.p, dsts, sizeof(int32_t) * n_edges, hipMemcpyHostToDevice));
  }
  return graph_from_device_edges(n_vertices, n_edges, s.p, d.p, out);
}

extern "C" int gr_graph_rmat(int scale, int edge_factor, double a, double b, double c,
                             uint64_t seed, int symmetrize, gr_graph** out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return set_error(GR_ERR_INVALID, "gr_graph_rmat: bad arguments");
  const int64_t m = (int64_t)edge_factor << scale;
  const int64_t mm = symmetrize ? 2 * m : m;
  if (mm >= ((int64_t)1 << 31)) return set_error(GR_ERR_INVALID, "gr_graph_rmat: too many edges");
  const RmatThresholds t = rmat_thresholds(a, b, c);
  const uint64_t base = stream_base(seed, kStreamRmat);
  int64_t n = 0;
  DevBuf<uint64_t> uniq;
  {
    DevBuf<uint64_t> raw, sorted;
    GR_TRY(raw.alloc(mm));
    GR_TRY(sorted.alloc(mm));
    k_rmat_keys<<<grid_for(m), kBlock>>>(base, m, scale, t, symmetrize, raw.p);
    GR_HIP(hipGetLastError());
    size_t temp = 0;
    GR_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, temp, raw.p, sorted.p, (int)mm, 0, 2 * scale));
    {
      DevBuf<uint8_t> tb;
      GR_TRY(tb.alloc((int64_t)temp));
      GR_HIP(hipcub::DeviceRadixSort::SortKeys(tb.p, temp, raw.p, sorted.p, (int)mm, 0, 2 * scale));
    }
    DevBuf<uint8_t> keep;
    GR_TRY(keep.alloc(mm));
    k_mark_unique<<<grid_for(mm), kBlock>>>(sorted.p, mm, scale, keep.p);
    GR_HIP(hipGetLastError());
    GR_TRY(uniq.alloc(mm));
    DevBuf<int> d_n;
    GR_TRY(d_n.alloc(1));
    temp = 0;
    GR_HIP(hipcub::DeviceSelect::Flagged(nullptr, temp, sorted.p, keep.p, uniq.p, d_n.p, (int)mm));
    {
      DevBuf<uint8_t> tb;
      GR_TRY(tb.alloc((int64_t)temp));
      GR_HIP(hipcub::DeviceSelect::Flagged(tb.p, temp, sorted.p, keep.p, uniq.p, d_n.p, (int)mm));
    }
    int hn = 0;
    GR_HIP(hipMemcpy(&hn, d_n.p, sizeof(int), hipMemcpyDeviceToHost));
    n = hn;
  }
  return graph_from_sorted_keys((int32_t)1 << scale, n, scale, uniq.p, out);
}

extern "C" int gr_graph_grid(int side, double delete_frac, uint64_t seed, gr_graph** out) {
  if (!out || side < 2 || side > 32768 || delete_frac < 0 || delete_frac > 1)
    return set_error(GR_ERR_INVALID, "gr_graph_grid: bad arguments");
  int64_t n = host_grid_arcs(side, delete_frac, seed, nullptr, nullptr);
  std::vector<int32_t> s((size_t)n), d((size_t)n);
  host_grid_arcs(side, delete_frac, seed, s.data(), d.data());
  return gr_graph_from_edges(side * side, (int32_t)n, s.data(), d.data(), out);
}

extern "C" int gr_graph_destroy(gr_graph* g) { return graph_free(g); }

extern "C" int gr_graph_info(const gr_graph* g, int32_t* n_vertices, int64_t* n_edges) {
  if (!g) return set_error(GR_ERR_INVALID, "graph is NULL");
  if (n_vertices) *n_vertices = g->nV;
  if (n_edges) *n_edges = g->nE;
  return GR_OK;
}

extern "C" int gr_graph_copy_csr(const gr_graph* g, int32_t* offsets, int32_t* dsts) {
  if (!g) return set_error(GR_ERR_INVALID, "graph is NULL");
  if (offsets) memcpy(offsets, g->h_off_out.data(), sizeof(int32_t) * (g->nV + 1));
  if (dsts && g->nE) GR_HIP(hipMemcpy(dsts, g->dsts, sizeof(int32_t) * g->nE, hipMemcpyDeviceToHost));
  return GR_OK;
}

extern "C" int gr_graph_copy_csc(const gr_graph* g, int32_t* offsets, int32_t* srcs) {
  if (!g) return set_error(GR_ERR_INVALID, "graph is NULL");
  if (offsets) memcpy(offsets, g->h_off_in.data(), sizeof(int32_t) * (g->nV + 1));
  if (srcs && g->nE) GR_HIP(hipMemcpy(srcs, g->srcs, sizeof(int32_t) * g->nE, hipMemcpyDeviceToHost));
  return GR_OK;
}

extern "C" int gr_graph_copy_edges(const gr_graph* g, int32_t* srcs, int32_t* dsts) {
  if (!g || !srcs || !dsts) return set_error(GR_ERR_INVALID, "bad arguments");
  const int64_t n = g->nE;
  std::vector<int32_t> adj((size_t)n), eid;
  if (n) GR_HIP(hipMemcpy(adj.data(), g->dsts, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  if (g->csr_eid) {
    eid.resize((size_t)n);
    if (n) GR_HIP(hipMemcpy(eid.data(), g->csr_eid, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  }
  for (int32_t v = 0; v < g->nV; ++v) {
    for (int32_t p = g->h_off_out[v]; p < g->h_off_out[v + 1]; ++p) {
      int32_t e = g->csr_eid ? eid[p] : p;
      srcs[e] = v;
      dsts[e] = adj[p];
    }
  }
  return GR_OK;
}

extern "C" int gr_graph_copy_perm(const gr_graph* g, int32_t* csr_eid, int32_t* csc_eid) {
  if (!g) return set_error(GR_ERR_INVALID, "graph is NULL");
  const int64_t n = g->nE;
  if (csr_eid) {
    if (g->csr_eid) {
      if (n) GR_HIP(hipMemcpy(csr_eid, g->csr_eid, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    } else {
      for (int64_t i = 0; i < n; ++i) csr_eid[i] = (int32_t)i;
    }
  }
  if (csc_eid && n)
    GR_HIP(hipMemcpy(csc_eid, g->csc_eid, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  return GR_OK;
}

extern "C" int gr_graph_max_out_degree_vertex(const gr_graph* g, int32_t* vertex, int32_t* degree) {
  if (!g || !vertex) return set_error(GR_ERR_INVALID, "bad arguments");
  int32_t best = -1, bestDeg = -1;
  for (int32_t v = 0; v < g->nV; ++v) {
    int32_t d = g->h_off_out[v + 1] - g->h_off_out[v];
    if (d > bestDeg) {  // strict: first vertex wins ties (bfs.cu:187)
      bestDeg = d;
      best = v;
    }
  }
  *vertex = best;
  if (degree) *degree = bestDeg;
  return GR_OK;
}

extern "C" int gr_graph_hash_weights_device(const gr_graph* g, uint64_t seed, int32_t lo,
                                            int span_log2, int32_t** device_weights) {
  if (!g || !device_weights || span_log2 < 1 || span_log2 > 31)
    return set_error(GR_ERR_INVALID, "bad arguments");
  DevBuf<int32_t> w;
  GR_TRY(w.alloc(g->nE));
  if (g->nE) {
    k_hash_weights<<<grid_for(g->nE), kBlock>>>(stream_base(seed, kStreamWeight), g->nE,
                                                (uint32_t)lo, span_log2, w.p);
    GR_HIP(hipGetLastError());
    GR_HIP(hipDeviceSynchronize());
  }
  *device_weights = w.release();
  return GR_OK;
}
