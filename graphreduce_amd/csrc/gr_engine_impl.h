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
//                    k_pull_stream / k_pull_heavy test each vertex's in-edges
//                    against the changed bitmap.
//   nextIter         k_compact_count/scan: the frontier bitmap becomes the
//                    active bitmap, counted with its in-degrees; the counts
//                    reach the host through mapped memory.  k_compact_write
//                    turns the active bitmap into the ascending active list
//                    plus the exclusive scan of its in-degrees (the gather's
//                    row offsets) when a phase needs the list; apply runs
//                    from the bitmap otherwise.
//
// This header holds the engine's kernels and the Engine<Program> template; it
// is shared by the library (which instantiates the built-in programs in
// gr_engine.hip) and by code that instantiates the engine for a program of
// its own through include/graphreduce_amd/gas_program.hpp.  Kernels that are
// not templates are static: every translation unit keeps its own copy.
#ifndef GRAPHREDUCE_AMD_GR_ENGINE_IMPL_H
#define GRAPHREDUCE_AMD_GR_ENGINE_IMPL_H

#include <string.h>
#include <algorithm>
#include <chrono>
#include <type_traits>
#include <vector>
#include "gr_device.h"
#include "gr_gather.h"

namespace gr {

// ------------------------------------------------------------------- apply

// Sets bit v of `bits` for the lanes with set != 0.  The lanes hold
// ascending vertex ids (lanes without a vertex pass valid = false and come
// last), so lanes sharing a 32-bit word are contiguous: a segmented OR over
// the wave leaves one atomic per word instead of one per vertex.
__device__ __forceinline__ void wave_set_bits_sorted(bool valid, bool set, int v,
                                                     uint32_t* __restrict__ bits) {
  if (!__ballot(set)) return;
  const int lane = lane_id();
  const uint32_t w = valid ? ((uint32_t)v >> 5) : 0xFFFFFFFFu;
  uint32_t b = set ? (1u << (v & 31)) : 0u;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t ow = __shfl_down(w, off, kWave);
    const uint32_t ob = __shfl_down(b, off, kWave);
    if (lane + off < kWave && ow == w) b |= ob;
  }
  const uint32_t pw = __shfl_up(w, 1, kWave);
  const bool leader = (lane == 0) || pw != w;
  if (leader && b && w != 0xFFFFFFFFu) atomicOr(&bits[w], b);
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

constexpr int kApplyNT = 256;
constexpr int kApplyVT = 8;
constexpr int kApplyChunk = kApplyNT * kApplyVT;

// Deferred changed list of a bitmap apply.  cc[c]: chunk c's (count << 32 |
// out-edges) of the vertices whose apply returned true and that have
// out-edges.  Chunks form groups of kApplyGroup; each chunk adds ONE word to
// its group's sc[group * kGroupStride] (groups 128 bytes apart): the chunk
// count in bits 50 and up, its vertex count in bits 32-49 (at most
// kApplyGroup * kApplyChunk = 2^17 per group), its out-edges in bits 0-31
// (the packed totals keep out-edges below 2^32, as ctr->chg does).  The
// chunk whose add completes the group knows the group total from the value
// the add returned, keeps it in sc_final[group], resets sc and adds the
// total to ctr->chg: one same-address atomic per group instead of one per
// chunk, and no fence.  k_chg_write finds a chunk's list position from
// sc_final and cc.
constexpr int kApplyGroup = 64;
constexpr int kGroupStride = 16;  // unsigned long longs between groups
constexpr int kGroupDoneShift = 50;
static_assert(kApplyGroup * kApplyChunk < (1 << (kGroupDoneShift - 32)),
              "a group's vertex count fits below the done count");
struct ApplyDefer {
  unsigned long long* cc = nullptr;  // nullptr: the list is appended at once
  unsigned long long* sc = nullptr;
  unsigned long long* sc_final = nullptr;
};

__device__ __forceinline__ void defer_chunk_done(const ApplyDefer& d, int c, int n_chunks,
                                                 unsigned long long tot,
                                                 DevCounters* __restrict__ ctr) {
  d.cc[c] = tot;
  const int grp = c / kApplyGroup;
  const unsigned members = (unsigned)min(kApplyGroup, n_chunks - grp * kApplyGroup);
  const unsigned long long mine = (1ull << kGroupDoneShift) + tot;
  const unsigned long long old = atomicAdd(&d.sc[grp * kGroupStride], mine);
  if ((unsigned)(old >> kGroupDoneShift) == members - 1) {
    const unsigned long long g = (old + mine) & ((1ull << kGroupDoneShift) - 1);
    d.sc[grp * kGroupStride] = 0ull;  // every chunk of the group has added
    d.sc_final[grp] = g;
    if (g) atomicAdd(&ctr->chg, g);
  }
}

// A one-byte degree of kDegSat stands for kDegSat or more: the offsets hold
// the degree.
constexpr int kDegSat = 255;

// The kept source value of an apply's entry.  Without flag_prev it is
// rewritten where the vertex data changed.  With flag_prev (programs with
// source_flag_v, bitmap form): its sign bit is set where the apply returned
// true with out-edges (`want`) and cleared where the previous apply set it
// (flag_prev, that apply's changed bitmap) and this one does not, so that
// the flags are exactly this apply's changed set afterwards.
template <class P, bool kBits>
__device__ __forceinline__ void keep_source_value(source_value_t<P>* __restrict__ srcval,
                                                  const uint32_t* __restrict__ flag_prev,
                                                  const uint32_t* s_pwords, int e, int n, int chunk,
                                                  int v, const typename P::VertexData& vs,
                                                  bool valid, bool changed, bool want) {
  if constexpr (source_value_v<P>) {
    if constexpr (source_flag_v<P> && kBits) {
      if (flag_prev) {
        const int i = chunk + e;
        const bool was = i < n && ((s_pwords[e >> 5] >> (i & 31)) & 1u);
        if (valid && (changed || want || was)) {
          const float sv = P::sourceValue(vs);
          srcval[v] = want ? flagged(sv) : sv;
        } else if (was) {
          srcval[v] = unflag(srcval[v]);
        }
        return;
      }
    }
    if (changed) srcval[v] = P::sourceValue(vs);
  }
}

// Program::apply over the active set: the ascending active list (kBits
// false: entry i is vertex active[i]) or the active bitmap (kBits true: entry
// i is vertex i, taken when its bit is set; n = nV).  Each workgroup takes a
// chunk of kApplyChunk entries, thread t entries chunk + m * kApplyNT + t.
// The vertices whose apply returned true and that have out-edges are appended
// to the changed list with ONE atomic per chunk, in entry order (ascending
// vertex ids within the chunk), so that a wave's lanes store to consecutive
// positions, with their out-edge offsets increasing with the list index;
// their bits are set in the changed bitmap for pull mode.
// Deferred form (defer.cc != nullptr, bitmap form with a changed bitmap):
// no list is written; each chunk's counts go to ApplyDefer (above), from
// which k_chg_write makes the list when a phase needs it.  One same-address
// atomic per workgroup serialises at the memory side (about 12 ns each; 8192
// of them took 100 us in one launch on an MI355X), which is what the
// deferred form avoids for large active sets.
template <class P, bool kBits>
__global__ __launch_bounds__(kApplyNT) void k_apply(
    int n, const int32_t* __restrict__ active, const uint32_t* __restrict__ act_bits,
    const typename P::GatherResult* __restrict__ gtmp, int use_gather,
    typename P::VertexData* __restrict__ vd, const int32_t* __restrict__ off_out,
    const uint8_t* __restrict__ deg8, int iteration, int32_t* __restrict__ chg_v,
    int32_t* __restrict__ chg_e,
    uint32_t* __restrict__ chg_bits, DevCounters* __restrict__ ctr,
    source_value_t<P>* __restrict__ srcval, typename P::GatherResult* __restrict__ cand,
    unsigned long long* __restrict__ applied, ApplyDefer defer,
    const uint32_t* __restrict__ flag_prev = nullptr, int gate_fab = 0) {
  // gate_fab: queued behind a blocked gather launched in place of the pull;
  // it runs only where that gather ran (ctr->fab, the same for every thread)
  if (gate_fab && ctr->fab == 0) return;
  constexpr int kNW = kApplyNT / kWave;
  constexpr int kWords = kApplyChunk / 32;
  static_assert(kWords <= kApplyNT, "one bitmap word per thread");
  __shared__ uint32_t s_words[kWords];
  __shared__ uint32_t s_pwords[kWords];  // flag_prev words of the chunk
  __shared__ int s_cnt[kApplyVT * kNW], s_deg[kApplyVT * kNW];
  __shared__ unsigned long long s_base;
  __shared__ unsigned s_applied;
  const int tid = threadIdx.x, wv = tid / kWave;
  if (tid == 0) s_applied = 0;
  unsigned my_applied = 0;  // applies that returned true, this thread
  for (int chunk = blockIdx.x * kApplyChunk; chunk < n; chunk += gridDim.x * kApplyChunk) {
    if constexpr (kBits) {
      // the chunk's bitmap words; a chunk without active vertices is skipped
      // (uniform: every thread sees the same barrier result)
      uint32_t word = 0u, pword = 0u;
      if (tid < kWords) {
        const int wi = (chunk >> 5) + tid;
        word = wi * 32 < n ? act_bits[wi] : 0u;
        s_words[tid] = word;
        if (flag_prev) {
          pword = wi * 32 < n ? flag_prev[wi] : 0u;
          s_pwords[tid] = pword;
        }
      }
      if (!__syncthreads_or((word | pword) != 0u)) {
        if (defer.cc && tid == 0)
          defer_chunk_done(defer, chunk / kApplyChunk, (n + kApplyChunk - 1) / kApplyChunk, 0ull,
                           ctr);
        continue;
      }
    }
    // In phases over the lane's kApplyVT entries (ids and gather values,
    // vertex values, apply, out-degrees of the changed ones, then the stores
    // and bit atomics), so that the loads of all entries are in flight
    // together: a store or atomic between them would make later waits for
    // loaded data wait for it too.
    typedef typename P::VertexData VD;
    int dvs[kApplyVT], degs[kApplyVT];
    typename P::GatherResult gs[kApplyVT];
    VD vs[kApplyVT];
    unsigned valid_mask = 0, r_mask = 0, chg_mask = 0, want_mask = 0;
#pragma unroll
    for (int m = 0; m < kApplyVT; ++m) {
      const int i = chunk + m * kApplyNT + tid;
      dvs[m] = 0;
      gs[m] = P::gatherZero;
      if constexpr (kBits) {
        dvs[m] = i;
        if (i < n && ((s_words[(m * kApplyNT + tid) >> 5] >> (i & 31)) & 1u))
          valid_mask |= 1u << m;
      } else if (i < n) {
        valid_mask |= 1u << m;
        dvs[m] = active[i];
        if (use_gather == 1) gs[m] = gtmp[i];
      }
    }
#pragma unroll
    for (int m = 0; m < kApplyVT; ++m)
      if (valid_mask & (1u << m)) vs[m] = vd[dvs[m]];
    // bitmap form: one-byte out-degrees of every entry, loaded with the
    // vertex values (not after apply: that would add a round trip to memory;
    // the entries are consecutive vertices, so the bytes share lines); vertex
    // 0's for the others keeps the loads unconditional
    uint32_t d8[kApplyVT];
    if constexpr (kBits) {
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m) d8[m] = deg8[(valid_mask & (1u << m)) ? dvs[m] : 0];
    }
    // use_gather 2 (push gathers of delta programs): the relaxed candidate of
    // each active row is its gather value; read it and reset it to gatherZero
    if (use_gather == 2) {
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m)
        if (valid_mask & (1u << m)) gs[m] = cand[dvs[m]];
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m)
        if (valid_mask & (1u << m)) cand[dvs[m]] = P::gatherZero;
    }
    // use_gather 3 (blocked gather): cand holds every vertex's gather value
    if (use_gather == 3) {
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m)
        if (valid_mask & (1u << m)) gs[m] = cand[dvs[m]];
    }
#pragma unroll
    for (int m = 0; m < kApplyVT; ++m) {
      if (valid_mask & (1u << m)) {
        const VD old = vs[m];
        if (P::apply(&vs[m], gs[m], iteration)) r_mask |= 1u << m;
        if (!pod_equal(vs[m], old)) chg_mask |= 1u << m;
      }
    }
    // out-degrees of the entries whose apply returned true (list form: only
    // theirs are loaded, the entries being scattered vertices); the offsets
    // only for a degree of kDegSat or more
    if constexpr (!kBits) {
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m) d8[m] = deg8[(r_mask & (1u << m)) ? dvs[m] : 0];
    }
#pragma unroll
    for (int m = 0; m < kApplyVT; ++m) {
      degs[m] = (r_mask & (1u << m)) ? (int)d8[m] : 0;
      if (degs[m] == kDegSat) degs[m] = off_out[dvs[m] + 1] - off_out[dvs[m]];
    }
    // positions in entry order: per (m, wave) the count of changed entries
    // and the sum of their out-degrees, scanned below
    unsigned long long ball[kApplyVT];
    int incl[kApplyVT];
    if (defer.cc) {  // deferred: stores, bits and the chunk's counts only
      unsigned long long mine = 0ull;
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m) {
        const bool want = degs[m] > 0;
        if (chg_mask & (1u << m)) vd[dvs[m]] = vs[m];
        keep_source_value<P, kBits>(srcval, flag_prev, s_pwords, m * kApplyNT + tid, n, chunk,
                                    dvs[m], vs[m], (valid_mask >> m) & 1u, (chg_mask >> m) & 1u,
                                    want);
        wave_set_bits_sorted(chunk + m * kApplyNT + tid < n, want, dvs[m], chg_bits);
        if (want) mine += (1ull << 32) | (unsigned long long)(uint32_t)degs[m];
      }
      my_applied += __popc(r_mask);
      mine = wave_reduce_add(mine);
      if (tid == 0) s_base = 0ull;
      __syncthreads();
      if (lane_id() == 0 && mine) atomicAdd(&s_base, mine);
      __syncthreads();
      if (tid == 0)
        defer_chunk_done(defer, chunk / kApplyChunk, (n + kApplyChunk - 1) / kApplyChunk, s_base,
                         ctr);
      __syncthreads();  // the next chunk reuses s_words and s_base
      continue;
    }
#pragma unroll
    for (int m = 0; m < kApplyVT; ++m) {
      const bool want = degs[m] > 0;  // apply returned true and out-edges exist
      if (chg_mask & (1u << m)) vd[dvs[m]] = vs[m];
      keep_source_value<P, kBits>(srcval, flag_prev, s_pwords, m * kApplyNT + tid, n, chunk,
                                  dvs[m], vs[m], (valid_mask >> m) & 1u, (chg_mask >> m) & 1u,
                                  want);
      if (chg_bits) wave_set_bits_sorted(chunk + m * kApplyNT + tid < n, want, dvs[m], chg_bits);
      if (want) want_mask |= 1u << m;
      ball[m] = __ballot(want);
      incl[m] = wave_incl_scan_add(want ? degs[m] : 0);
      if (lane_id() == kWave - 1) {
        s_cnt[m * kNW + wv] = __popcll(ball[m]);
        s_deg[m * kNW + wv] = incl[m];
      }
    }
    my_applied += __popc(r_mask);
    __syncthreads();
    if (wv == 0) {  // the (m, wave) table's exclusive scan, one entry per lane
      static_assert(kApplyVT * kNW <= kWave, "the table fits one wave");
      const int lane = lane_id();
      const int a = lane < kApplyVT * kNW ? s_cnt[lane] : 0;
      const int b = lane < kApplyVT * kNW ? s_deg[lane] : 0;
      const int ia = wave_incl_scan_add(a), ib = wave_incl_scan_add(b);
      if (lane < kApplyVT * kNW) {
        s_cnt[lane] = ia - a;
        s_deg[lane] = ib - b;
      }
      const int c = __shfl(ia, kWave - 1, kWave), e = __shfl(ib, kWave - 1, kWave);
      if (lane == 0)
        s_base = c ? atomicAdd(&ctr->chg, ((unsigned long long)c << 32) |
                                              (unsigned long long)(uint32_t)e)
                   : 0ull;
    }
    __syncthreads();
    if (want_mask) {
      const unsigned long long base = s_base;
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m) {
        if (want_mask & (1u << m)) {
          const int idx = (int)(base >> 32) + s_cnt[m * kNW + wv] + lanes_below(ball[m]);
          chg_v[idx] = dvs[m];
          chg_e[idx] = (int)(uint32_t)base + s_deg[m * kNW + wv] + incl[m] - degs[m];
        }
      }
    }
    __syncthreads();  // the next chunk reuses s_words, s_cnt, s_deg and s_base
  }
  // every apply that returned true counts toward sum #C, also those of
  // vertices without out-edges, which join no changed list
  const unsigned wsum = wave_reduce_add(my_applied);
  __syncthreads();
  if (lane_id() == 0 && wsum) atomicAdd(&s_applied, wsum);
  __syncthreads();
  if (tid == 0 && s_applied)
    atomicAdd(&applied[(blockIdx.x % kAppliedSlots) * kAppliedStride],
              (unsigned long long)s_applied);
}

// The changed list of a deferred apply: chunk c (kApplyChunk vertices, the
// 64 words of the changed bitmap from word 64 c) starts at the list index and
// out-edge offset that the groups before its group (sc_final) and the chunks
// of its group before it (cc) add up to; thread t takes byte t of the chunk
// (8 vertices), the list is ascending in vertex id, each entry's out-edge
// offset the running sum of the out-degrees before it.  A chunk's entries are
// staged in LDS and written out coalesced.
constexpr int kChgNT = kApplyChunk / 8;
//
// standdown != 0: when scatterActivate's pull (or the blocked gather in its
// place) runs, by the same rule on the same counters, no phase of this
// iteration or the next one reads the list: nothing is written then.
[[maybe_unused]] static __global__ __launch_bounds__(kChgNT) void k_chg_write(
    const uint32_t* __restrict__ chg_bits, int nV, const int32_t* __restrict__ off_out,
    const uint8_t* __restrict__ deg8, const unsigned long long* __restrict__ cc,
    const unsigned long long* __restrict__ sc_final,
    int n_chunks, int32_t* __restrict__ chg_v, int32_t* __restrict__ chg_e,
    const DevCounters* __restrict__ ctr, int standdown, int mode, long long threshold,
    long long fuse_max) {
  __shared__ unsigned long long s_scan[kChgNT / kWave];
  __shared__ unsigned long long s_pref;
  __shared__ int32_t s_v[kApplyChunk], s_e[kApplyChunk];
  if (standdown) {  // uniform: every workgroup reads the same counters
    const unsigned long long c = ctr->chg;
    if (use_pull_rule(c, mode, threshold) && !fused_scatter_rule(c, fuse_max)) return;
  }
  for (int c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const unsigned long long own = cc[c];
    if (own == 0ull) continue;  // uniform: no entries in this chunk
    if (threadIdx.x < kWave) {
      const int grp = c / kApplyGroup;
      unsigned long long p = 0ull;
      for (int i = threadIdx.x; i < grp; i += kWave) p += sc_final[i];
      for (int i = grp * kApplyGroup + (int)threadIdx.x; i < c; i += kWave) p += cc[i];
      p = wave_reduce_add(p);
      if (threadIdx.x == 0) s_pref = p;
    }
    __syncthreads();
    const unsigned long long b0 = s_pref;
    const int v0 = c * kApplyChunk + (int)threadIdx.x * 8;
    const uint32_t word = v0 < nV ? chg_bits[v0 >> 5] : 0u;
    uint32_t bits = (word >> ((threadIdx.x & 3) * 8)) & 0xFFu;
    int dg[8];
    unsigned long long mine = 0ull;
    if (bits) {
      // one-byte out-degrees of the 8 vertices (v0 < nV, a multiple of 8:
      // deg8 covers them); the offsets for a degree of kDegSat or more
      const uint2 d = *reinterpret_cast<const uint2*>(deg8 + v0);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        dg[j] = (int)(((j < 4 ? d.x : d.y) >> (8 * (j & 3))) & 0xFFu);
        if (((bits >> j) & 1u) && dg[j] == kDegSat) dg[j] = off_out[v0 + j + 1] - off_out[v0 + j];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if ((bits >> j) & 1u) mine += (1ull << 32) | (unsigned long long)(uint32_t)dg[j];
    }
    unsigned long long tot;
    const unsigned long long ex = block_excl_scan_add<kChgNT>(mine, s_scan, &tot);
    int pos = (int)(ex >> 32);
    int dpos = (int)(uint32_t)b0 + (int)(uint32_t)ex;
    while (bits) {
      const int j = __builtin_ctz(bits);
      bits &= bits - 1;
      s_v[pos] = v0 + j;
      s_e[pos] = dpos;
      dpos += dg[j];
      ++pos;
    }
    __syncthreads();
    const int cnt = (int)(tot >> 32), first = (int)(b0 >> 32);
    for (int i = threadIdx.x; i < cnt; i += kChgNT) {
      chg_v[first + i] = s_v[i];
      chg_e[first + i] = s_e[i];
    }
    __syncthreads();  // the next chunk stages into the same arrays
  }
}

// srcval[v] = P::sourceValue(vd[v]) for every vertex.
template <class P>
__global__ void k_source_values(int nV, const typename P::VertexData* __restrict__ vd,
                                source_value_t<P>* __restrict__ srcval) {
  if constexpr (source_value_v<P>) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x)
      srcval[v] = P::sourceValue(vd[v]);
  }
}

// Vertex scan for a gather without a previous changed set, for programs with
// inert sources: writes the bitmap of the non-inert vertices (every word < 
// n_words, so nothing needs clearing first) and appends the non-inert ones
// with out-edges to the changed list as k_apply does (one atomic per chunk,
// out-edge offsets increasing with the list index), counted in ctr->inert.
template <class P>
__global__ __launch_bounds__(kApplyNT) void k_inert_scan(
    int nV, int n_words, const typename P::VertexData* __restrict__ vd,
    const int32_t* __restrict__ off_out, int32_t* __restrict__ chg_v, int32_t* __restrict__ chg_e,
    uint32_t* __restrict__ bits, DevCounters* __restrict__ ctr) {
  __shared__ int s_a[kApplyNT / kWave], s_b[kApplyNT / kWave];
  __shared__ unsigned long long s_base;
  const int tid = threadIdx.x, lane = lane_id();
  for (int chunk = blockIdx.x * kApplyChunk; chunk < nV; chunk += gridDim.x * kApplyChunk) {
    int degs[kApplyVT];
    unsigned want_mask = 0;
    int my_cnt = 0, my_edges = 0;
#pragma unroll
    for (int m = 0; m < kApplyVT; ++m) {
      // a wave covers the 64 consecutive vertices from v - lane (a multiple of 64)
      const int v = chunk + m * kApplyNT + tid;
      bool live = false;
      int deg = 0;
      if (v < nV) {
        live = !P::inertSource(vd[v]);
        if (live) deg = off_out[v + 1] - off_out[v];
      }
      const uint64_t word2 = __ballot(live);
      const int w0 = (v - lane) >> 5;
      if (lane == 0 && w0 < n_words) bits[w0] = (uint32_t)word2;
      if (lane == 32 && w0 + 1 < n_words) bits[w0 + 1] = (uint32_t)(word2 >> 32);
      degs[m] = deg;
      if (live && deg > 0) {
        want_mask |= 1u << m;
        my_cnt += 1;
        my_edges += deg;
      }
    }
    int tc, te;
    const int ec = block_excl_scan_add<kApplyNT>(my_cnt, s_a, &tc);
    const int ee = block_excl_scan_add<kApplyNT>(my_edges, s_b, &te);
    if (tc > 0) {
      if (tid == 0)
        s_base = atomicAdd(&ctr->inert, ((unsigned long long)tc << 32) |
                                            (unsigned long long)(uint32_t)te);
      __syncthreads();
      const unsigned long long base = s_base;
      int idx = (int)(base >> 32) + ec;
      int eo = (int)(uint32_t)base + ee;
#pragma unroll
      for (int m = 0; m < kApplyVT; ++m) {
        if (want_mask & (1u << m)) {
          chg_v[idx] = chunk + m * kApplyNT + tid;
          chg_e[idx] = eo;
          ++idx;
          eo += degs[m];
        }
      }
      __syncthreads();
    }
  }
}

// *bad != 0 when some edge's data fails P::edgeKeepsInert.
template <class P>
__global__ void k_check_inert_edges(int64_t E, const typename P::EdgeData* __restrict__ ed,
                                    unsigned int* __restrict__ bad) {
  bool b = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x)
    b = b || !P::edgeKeepsInert(ed[i]);
  if (__ballot(b) && lane_id() == 0) atomicOr(bad, 1u);
}

// ------------------------------------------------------- scatter: push

// Direction of scatterActivate, decided on the device from the changed
// out-edge count so that the host needs no read: mode 0 = pull when the
// out-edges of the changed vertices exceed `threshold`, 1 = push, 2 = pull.
__device__ __forceinline__ bool use_pull(unsigned long long chg, int mode, long long threshold) {
  return use_pull_rule(chg, mode, threshold);
}

// Fused activation + push gather (delta-gather programs): when the changed
// vertices' out-edges number at most fuse_max (>= 0), ONE pass over them sets
// the frontier bits and lowers the next gather's candidates, replacing the
// pull or push activation and the next iteration's push relax.
__device__ __forceinline__ bool fused_scatter(unsigned long long chg, long long fuse_max) {
  return fused_scatter_rule(chg, fuse_max);
}

// A fused pass over more than kScanMinEdges out-edges leaves the marking of
// most targets to k_cand_to_bits: a target whose candidate ends below
// gatherZero was relaxed, so one sequential scan of the candidates replaces a
// random bitmap update per edge.
constexpr long long kScanMinEdges = 1 << 20;
__device__ __forceinline__ bool fused_scan(unsigned long long chg, long long fuse_max) {
  return fused_scatter(chg, fuse_max) && (long long)(uint32_t)chg > kScanMinEdges;
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
// (also for the fused pass; write_fused: record in ctr->fused whether it runs)
[[maybe_unused]] static __global__ void k_push_partition(DevCounters* __restrict__ ctr, int mode, long long threshold,
                                 const int32_t* __restrict__ chg_e, int32_t* __restrict__ part,
                                 int max_tiles, long long fuse_max, int write_fused) {
  const unsigned long long c = ctr->chg;
  const bool fused = fused_scatter(c, fuse_max);
  if (write_fused && blockIdx.x == 0 && threadIdx.x == 0) ctr->fused = fused ? 1 : 0;
  if (use_pull(c, mode, threshold) && !fused) return;
  const int n_c = (int)(c >> 32);
  const int w_c = (int)(uint32_t)c;
  const int n_tiles = (w_c + kTile - 1) / kTile;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= n_tiles && t <= max_tiles;
       t += gridDim.x * blockDim.x) {
    int e = min(t * kTile, w_c);
    part[t] = max(0, upper_bound_dev(chg_e, n_c, e) - 1);
  }
}

// Row of each of a push tile's n_e edges (index into the tile's changed-list
// entries s_start[0..nr)): each lane searches the row of its first of kVT
// consecutive edges and walks the rest.  The changed list holds only vertices
// with out-edges, so the starts rise strictly and nr <= kTile + 1.  The
// callers then read the row of a strided edge from s_row instead of searching
// for it.
__device__ __forceinline__ void label_push_rows(const int32_t* s_start, int nr, int e0, int n_e,
                                                uint16_t* s_row) {
  const int l0 = threadIdx.x * kVT;
  if (l0 >= n_e) return;
  const int e = e0 + l0;
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (s_start[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  int next = lo + 1 < nr ? s_start[lo + 1] : INT_MAX;
  for (int q = 0; q < kVT && l0 + q < n_e; ++q) {
    while (next <= e + q) {
      ++lo;
      next = lo + 1 < nr ? s_start[lo + 1] : INT_MAX;
    }
    s_row[l0 + q] = (uint16_t)lo;
  }
}

// Push activation.  Targets below n_hub (the highest-degree vertices in the
// degree-ordered layout) are marked with a plain byte store in hub_flags,
// which nextIter's compaction folds into the bitmap: many changed vertices share a few
// hub targets, and atomics on the same word are serialised where they are
// performed, while stores to one line merge in each L2.  The other targets
// are marked with an atomicOr unless their bit is already visible.
[[maybe_unused]] static __global__ __launch_bounds__(kNT) void k_push_tiles(const DevCounters* __restrict__ ctr,
                                                    int mode, long long threshold,
                                                    long long fuse_max,
                                                    const int32_t* __restrict__ part,
                                                    const int32_t* __restrict__ chg_v,
                                                    const int32_t* __restrict__ chg_e,
                                                    const int32_t* __restrict__ off_out,
                                                    const int32_t* __restrict__ dsts,
                                                    uint32_t* __restrict__ bitmap,
                                                    uint8_t* __restrict__ hub_flags, int n_hub) {
  __shared__ int32_t s_start[kTile + 2];
  __shared__ int32_t s_csr[kTile + 2];
  __shared__ uint16_t s_row[kTile];
  const unsigned long long c = ctr->chg;
  if (use_pull(c, mode, threshold) || fused_scatter(c, fuse_max)) return;
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
    label_push_rows(s_start, nr, e0, e1 - e0, s_row);
    __syncthreads();
    // The lane's kVT targets, then their bitmap words, are all read before its
    // first atomic: the compiler keeps later loads behind an atomic that may
    // alias them, which would make the items one memory round trip each.
    int pos[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      const int e = e0 + tid + m * kNT;
      pos[m] = -1;
      if (e < e1) {
        const int r = s_row[e - e0];
        pos[m] = s_csr[r] + (e - s_start[r]);
      }
    }
    int dst[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m)
      dst[m] = pos[m] >= 0 ? __builtin_nontemporal_load(dsts + pos[m]) : -1;
    uint32_t word[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m) word[m] = dst[m] >= n_hub ? bitmap[dst[m] >> 5] : ~0u;
    // The decisions come first, so that no wait for loaded data (which would
    // also wait for the atomics already issued) falls between the atomics.
    uint32_t need = 0;
#pragma unroll
    for (int m = 0; m < kVT; ++m)
      if (dst[m] >= n_hub && !(word[m] & (1u << (dst[m] & 31)))) need |= 1u << m;
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      if (dst[m] >= 0 && dst[m] < n_hub) hub_flags[dst[m]] = 1;
      if (need & (1u << m)) atomicOr(&bitmap[dst[m] >> 5], 1u << (dst[m] & 31));
    }
    __syncthreads();
  }
}

// Program::scatter over the out-edges of the changed list (refgas.h:172-196):
// scatter(&vd[u], &vd[v], &edge) for every out-edge u -> v of a vertex whose
// apply returned true.  Vertex data is read only and every edge is written by
// its own source alone, so the calls are independent.  Edge data is in
// edge-list order, reached through csr_eid.
template <class P>
__global__ __launch_bounds__(kNT) void k_scatter_tiles(
    const DevCounters* __restrict__ ctr, const int32_t* __restrict__ part,
    const int32_t* __restrict__ chg_v, const int32_t* __restrict__ chg_e,
    const int32_t* __restrict__ off_out, const int32_t* __restrict__ dsts,
    const int32_t* __restrict__ csr_eid, const typename P::VertexData* __restrict__ vd,
    typename P::EdgeData* __restrict__ ed) {
  __shared__ int32_t s_start[kTile + 2];
  __shared__ int32_t s_csr[kTile + 2];
  __shared__ int32_t s_u[kTile + 2];
  const int w_c = (int)(uint32_t)ctr->chg;
  const int n_tiles = (w_c + kTile - 1) / kTile;
  const int tid = threadIdx.x;
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int k0 = part[t], k1 = part[t + 1];
    const int nr = k1 - k0 + 1;
    const int e0 = t * kTile, e1 = min(e0 + kTile, w_c);
    for (int k = tid; k < nr; k += kNT) {
      const int u = chg_v[k0 + k];
      s_u[k] = u;
      s_start[k] = chg_e[k0 + k];
      s_csr[k] = off_out[u];
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
        const int pos = s_csr[lo] + (e - s_start[lo]);
        P::scatter(vd + s_u[lo], vd + dsts[pos], ed + csr_eid[pos]);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------- delta gather by push
//
// For delta-gather programs the gather of the next iteration only needs the
// in-edges whose source changed, i.e. the out-edges of the changed list.  When
// those are few, they are pushed: every out-edge (u -> v) of a changed u
// lowers cand[v] to gatherMap(u, edge) with atomicMin, and the gather result of
// an active row is then cand of its vertex.  cand holds gatherZero everywhere
// between gathers: only out-neighbours of changed vertices receive values and
// they are exactly the next active list, whose entries are reset when read.

// part[t] = changed-list entry holding out-edge t * kTile (n_c entries, w_c edges)
[[maybe_unused]] static __global__ void k_relax_partition(int n_c, int w_c, const int32_t* __restrict__ chg_e,
                                  int32_t* __restrict__ part) {
  const int n_tiles = (w_c + kTile - 1) / kTile;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= n_tiles; t += gridDim.x * blockDim.x)
    part[t] = max(0, upper_bound_dev(chg_e, n_c, min(t * kTile, w_c)) - 1);
}

// Targets below kRelaxHub (the highest-degree vertices of the degree order)
// are folded per workgroup in LDS first: in the late iterations most changed
// vertices have few out-edges and those lead mostly to the same few hubs, whose
// atomics would be serialised where they are performed.  Each workgroup then
// lowers a hub's candidate (and marks it) once, after its last tile.
#ifndef GR_RELAX_HUB
#define GR_RELAX_HUB 2048
#endif
constexpr int kRelaxHub = GR_RELAX_HUB;
static_assert(kRelaxHub % kWave == 0, "hub folding flushes whole waves of ids");

template <class P>
// With ctr (the fused pass of scatterActivate): runs only when fused_scatter
// holds, takes the edge count from ctr->chg and also marks the targets in the
// frontier `bitmap`; with scan_marks above kScanMinEdges edges only those
// whose relaxed value is not below gatherZero (k_cand_to_bits finds the rest).
[[maybe_unused]] static __global__ __launch_bounds__(kNT) void k_relax_tiles(
    int w_c, const DevCounters* __restrict__ ctr, long long fuse_max,
    const int32_t* __restrict__ part, const int32_t* __restrict__ chg_v,
    const int32_t* __restrict__ chg_e, const int32_t* __restrict__ off_out,
    const int32_t* __restrict__ dsts, const typename P::VertexData* __restrict__ vd,
    const typename P::EdgeData* __restrict__ ed_csr, typename P::GatherResult* __restrict__ cand,
    uint32_t* __restrict__ bitmap, int scan_marks) {
  static_assert(!P::kGatherUsesDst, "push relax needs a gatherMap without the destination");
  typedef typename P::GatherResult G;
  constexpr int kHub = kRelaxHub > 0 ? kRelaxHub : kWave;
  __shared__ int32_t s_start[kTile + 2];
  __shared__ int32_t s_csr[kTile + 2];
  __shared__ typename P::VertexData s_vd[kTile + 2];
  __shared__ uint16_t s_row[kTile];
  __shared__ G s_hub[kHub];               // hub candidates of this workgroup
  __shared__ uint32_t s_hub_mark[kHub / 32];  // hubs to mark in the bitmap
  bool scan = false;
  if (ctr) {
    const unsigned long long c = ctr->chg;
    if (!fused_scatter(c, fuse_max)) return;
    w_c = (int)(uint32_t)c;
    scan = scan_marks && fused_scan(c, fuse_max);
  }
  const int n_tiles = (w_c + kTile - 1) / kTile;
  const int tid = threadIdx.x;
  if ((int)blockIdx.x >= n_tiles) return;  // uniform per workgroup
  if (kRelaxHub > 0) {
    for (int i = tid; i < kHub; i += kNT) s_hub[i] = P::gatherZero;
    for (int i = tid; i < kHub / 32; i += kNT) s_hub_mark[i] = 0u;
    // the first tile's __syncthreads orders these before any use
  }
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int k0 = part[t], k1 = part[t + 1];
    const int nr = k1 - k0 + 1;
    const int e0 = t * kTile, e1 = min(e0 + kTile, w_c);
    for (int k = tid; k < nr; k += kNT) {
      const int u = chg_v[k0 + k];
      s_start[k] = chg_e[k0 + k];
      s_csr[k] = off_out[u];
      s_vd[k] = vd[u];
    }
    __syncthreads();
    label_push_rows(s_start, nr, e0, e1 - e0, s_row);
    __syncthreads();
    // Targets and values of the lane's kVT edges, then every candidate (and
    // frontier word) read, then the atomics: the compiler keeps later loads
    // behind an atomic that may alias them, which would make the items one
    // memory round trip each.
    typedef typename P::GatherResult G;
    int pos[kVT], row[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      const int e = e0 + tid + m * kNT;
      pos[m] = -1;
      row[m] = 0;
      if (e < e1) {
        const int r = s_row[e - e0];
        pos[m] = s_csr[r] + (e - s_start[r]);
        row[m] = r;
      }
    }
    int vv[kVT];
    typename P::EdgeData ev[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      vv[m] = -1;
      if (pos[m] >= 0) {
        vv[m] = __builtin_nontemporal_load(dsts + pos[m]);
        ev[m] = stream_load(ed_csr + pos[m]);
      }
    }
    G val[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m)
      val[m] = pos[m] >= 0 ? P::gatherMap(nullptr, &s_vd[row[m]], &ev[m]) : P::gatherZero;
    if (kRelaxHub > 0) {
#pragma unroll
      for (int m = 0; m < kVT; ++m) {
        if (vv[m] >= 0 && vv[m] < kRelaxHub) {
          if (val[m] < P::gatherZero) atomicMin(&s_hub[vv[m]], val[m]);
          if (bitmap && !(scan && val[m] < P::gatherZero))
            atomicOr(&s_hub_mark[vv[m] >> 5], 1u << (vv[m] & 31));
          vv[m] = -1;  // done with this edge
        }
      }
    }
    G cur[kVT];
    uint32_t word[kVT];
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      cur[m] = vv[m] >= 0 ? cand[vv[m]] : P::gatherZero;
      const bool mark = vv[m] >= 0 && bitmap && !(scan && val[m] < P::gatherZero);
      word[m] = mark ? bitmap[vv[m] >> 5] : ~0u;
    }
    // decisions first: no wait for loaded data between the atomics
    uint32_t need_min = 0, need_or = 0;
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      if (vv[m] >= 0 && val[m] < cur[m]) need_min |= 1u << m;
      if (vv[m] >= 0 && !(word[m] & (1u << (vv[m] & 31)))) need_or |= 1u << m;
    }
#pragma unroll
    for (int m = 0; m < kVT; ++m) {
      if (need_min & (1u << m)) atomicMin(&cand[vv[m]], val[m]);
      if (need_or & (1u << m)) atomicOr(&bitmap[vv[m] >> 5], 1u << (vv[m] & 31));
    }
    __syncthreads();
  }
  if (kRelaxHub > 0) {
    // the workgroup's hub candidates and marks (the loop's last
    // __syncthreads made the LDS folds visible)
    for (int i = tid; i < kHub; i += kNT) {
      const G v = s_hub[i];
      if (v < P::gatherZero && v < cand[i]) atomicMin(&cand[i], v);
    }
    if (bitmap) {
      for (int k = tid; k < kHub / 32; k += kNT) {
        const uint32_t m = s_hub_mark[k];
        if (m && (bitmap[k] & m) != m) atomicOr(&bitmap[k], m);
      }
    }
  }
}

// After a fused pass with scan marks: bitmap |= {v : cand[v] < gatherZero}.
// The candidates were all gatherZero before the pass and only fall during it,
// and every relaxed value below gatherZero leaves its target below it, so the
// set is exactly those targets (the others were marked by the pass itself).
// One wave reads 64 consecutive candidates and writes their two bitmap words;
// cand holds whole 32-entry groups (the padding stays gatherZero).  Only this
// kernel writes the bitmap then: the other activation kernels stood down.
template <class P>
__global__ void k_cand_to_bits(const DevCounters* __restrict__ ctr, long long fuse_max,
                               int n_words, const typename P::GatherResult* __restrict__ cand,
                               uint32_t* __restrict__ bitmap) {
  if (!fused_scan(ctr->chg, fuse_max)) return;
  const long long total = (long long)n_words * 32;
  const int lane = threadIdx.x & 63;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total;
       base += stride) {
    const long long i = base + lane;
    const bool hit = i < total && cand[i] < P::gatherZero;
    const unsigned long long m = __ballot(hit);
    if (lane == 0 && m) {
      const long long w = base >> 5;
      if ((uint32_t)m) bitmap[w] |= (uint32_t)m;
      if ((m >> 32) && w + 1 < n_words) bitmap[w + 1] |= (uint32_t)(m >> 32);
    }
  }
}

// gtmp[i] = cand[active[i]] (already folded with gatherZero); resets cand.
template <class P>
__global__ void k_relax_collect(int n, const int32_t* __restrict__ active,
                                typename P::GatherResult* __restrict__ cand,
                                typename P::GatherResult* __restrict__ gtmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = active[i];
  gtmp[i] = cand[v];
  cand[v] = P::gatherZero;
}

template <class T>
__global__ void k_fill(int64_t n, T value, T* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = value;
}

// ------------------------------------------------------- scatter: pull

// next[v] = some in-neighbour of v is in the changed bitmap.
//
// k_pull_stream: one wave owns 64 consecutive vertices (lane l <-> vertex
// base + l), whose in-edges form one contiguous CSC segment.  The wave
// streams that segment 64 edges at a time (coalesced srcs reads); a ballot
// gives the hit lanes, and each lane tests whether its own vertex's slice of
// the chunk contains a hit (a second ballot), so no cross-lane shuffles are
// needed.  The row holding the chunk start is found from a ballot over the
// lanes' row offsets; rows already hit, and heavy rows (in-degree above
// kPullHeavy, done by k_pull_heavy with a wave per row in growing rounds and early exit),
// are jumped over.  Before that, each lane tests kPullFirst in-edges of its
// own row, heavy or not, the first and the last half of them, in two rounds
// from both ends inwards: rows list their sources by id, which the degree order
// makes hubs first, so a change among the hubs shows at the front of a row and
// one among the low-degree vertices at its back.  Rows with at most kPullFirst
// in-edges are settled either way, and the segment walk starts at the first
// open row; k_pull_heavy skips the heavy rows this test already marked.
#ifndef GR_PULL_HEAVY_FIRST
#define GR_PULL_HEAVY_FIRST 1
#endif
#ifndef GR_PULL_GRID
#define GR_PULL_GRID 4096
#endif
constexpr int kPullHeavy = 1024;
constexpr int kPullFirst = 4;
constexpr int kHubVertices = 1 << 20;  // push marks by byte store below this id

__device__ __forceinline__ uint64_t lane_range_mask(int lo, int hi) {  // bits [lo, hi)
  const uint64_t upto = hi >= 64 ? ~0ull : ((1ull << hi) - 1);
  const uint64_t below = lo <= 0 ? 0ull : ((1ull << lo) - 1);
  return upto & ~below;
}

[[maybe_unused]] static __global__ __launch_bounds__(256) void k_pull_stream(
    const DevCounters* __restrict__ ctr, int mode, long long threshold, long long fuse_max,
    int nV, const int32_t* __restrict__ off_in, const int32_t* __restrict__ srcs,
    const uint32_t* __restrict__ chg_bits, uint32_t* __restrict__ bitmap) {
  if (!use_pull(ctr->chg, mode, threshold) || fused_scatter(ctr->chg, fuse_max)) return;
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int nw = gridDim.x * wpb;
  const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (int)threadIdx.x / kWave);
  for (int base = w0 * kWave; base < nV; base += nw * kWave) {
    const int cnt = min(kWave, nV - base);
    const bool mine = lane < cnt;
    const int my_off = mine ? off_in[base + lane] : INT_MAX;
    const int my_end = mine ? off_in[base + lane + 1] : INT_MAX;
    const int deg = mine ? my_end - my_off : 0;
    const bool heavy = deg > kPullHeavy;
    // kPullFirst in-edges of the lane's own row, the first and the last half,
    // in two rounds from the ends inwards: the first and the last in-edge
    // settle most rows (at RMAT-24 the first alone 80 % of the rows with
    // in-edges in one pull, the last 62 % in the other), and the test's cost
    // is its random bitmap reads
    bool first_hit = false;
    if (deg > 0 && (GR_PULL_HEAVY_FIRST || !heavy)) {
#pragma unroll
      for (int r = 0; r < kPullFirst / 2; ++r) {
        const int pa = my_off + r, pz = my_end - 1 - r;  // from the front, from the back
        const int ua = !first_hit && pa < my_end ? srcs[pa] : -1;
        const int uz = !first_hit && pz > pa ? srcs[pz] : -1;
        const uint32_t wa = ua >= 0 ? chg_bits[ua >> 5] : 0u;
        const uint32_t wz = uz >= 0 ? chg_bits[uz >> 5] : 0u;
        first_hit |= (ua >= 0 && ((wa >> (ua & 31)) & 1u)) || (uz >= 0 && ((wz >> (uz & 31)) & 1u));
      }
    }
    uint64_t hits = __ballot(first_hit);
    // rows still open: more than kPullFirst in-edges, not heavy, no hit yet
    const uint64_t need = __ballot(deg > kPullFirst && !heavy && !first_hit);
    const uint64_t settled = ~need;  // for the walk: nothing left to find there
    // the walk's bounds are scalars (lane_value), so the loop is a scalar loop
    const int seg1 = lane_value(my_end, cnt - 1);
    int pb = seg1;
    if (need) pb = lane_value(my_off, __builtin_ctzll(need)) + kPullFirst / 2;
    while (pb < seg1 && (hits & need) != need) {
      // row holding edge pb: the last lane whose row starts at or before pb
      const uint64_t started = __ballot(mine && my_off <= pb);
      const int r0 = 63 - __builtin_clzll(started);
      if (((settled | hits) >> r0) & 1ull) {
        pb = lane_value(my_end, r0);
        continue;
      }
      const int p = pb + lane;
      bool h = false;
      if (p < seg1) {
        const int u = srcs[p];
        h = (chg_bits[u >> 5] >> (u & 31)) & 1u;
      }
      const uint64_t hit_lanes = __ballot(h);
      bool row_hit = false;
      if (mine && !heavy && hit_lanes) {
        const int lo = my_off - pb, hi = my_end - pb;
        if (hi > 0 && lo < kWave) row_hit = (hit_lanes & lane_range_mask(lo, hi)) != 0;
      }
      hits |= __ballot(row_hit);
      pb += kWave;
    }
    // lanes 0 and 32 write the two 32-bit words of this wave's vertices
    if ((lane & 31) == 0 && base + lane < nV) {
      const uint32_t word = (uint32_t)(hits >> lane);
      if (word) bitmap[(base + lane) >> 5] = word;
    }
  }
}

// The heavy rows' walk: one wave per row reads its in-edges in rounds of 64,
// 128, then 256 (up to four per lane, all in flight) until one comes from a
// changed vertex; rows that k_pull_stream's test of the first and last
// in-edges marked already are skipped.  The walk's cost is its random
// changed-bitmap reads: at RMAT-24 the heavy rows' first hits sit at a median
// of about 250 in-edges (at most about 2,500), so a wave reads little past
// the hit and the longest walk takes few rounds.  (A walk from both ends
// inwards was slower: 104 -> 122 us in BFS's dense pull.)  Row bounds are
// scalars.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_pull_heavy(
    const DevCounters* __restrict__ ctr, int mode, long long threshold, long long fuse_max,
    const int32_t* __restrict__ heavy, int n_heavy, const int32_t* __restrict__ off_in,
    const int32_t* __restrict__ srcs, const uint32_t* __restrict__ chg_bits,
    uint32_t* __restrict__ bitmap) {
  if (!use_pull(ctr->chg, mode, threshold) || fused_scatter(ctr->chg, fuse_max)) return;
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int nw = gridDim.x * wpb;
  const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (int)threadIdx.x / kWave);
  for (int h = w0; h < n_heavy; h += nw) {
    const int v = __builtin_amdgcn_readfirstlane(heavy[h]);
    // one load instruction: every lane of the wave sees the same word
    if ((__builtin_amdgcn_readfirstlane(bitmap[v >> 5]) >> (v & 31)) & 1u) continue;
    const int b = __builtin_amdgcn_readfirstlane(off_in[v]);
    const int e = __builtin_amdgcn_readfirstlane(off_in[v + 1]);
    int k = 1;  // 64-edge steps this round
    for (int pb = b; pb < e; pb += k * kWave, k = k < 4 ? 2 * k : 4) {
      bool hit = false;
      int u[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = pb + j * kWave + lane;
        u[j] = (j < k && p < e) ? srcs[p] : -1;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (u[j] >= 0) hit |= (chg_bits[u[j] >> 5] >> (u[j] & 31)) & 1u;
      if (__ballot(hit)) {
        if (lane == 0) atomicOr(&bitmap[v >> 5], 1u << (v & 31));
        break;
      }
    }
  }
}

// -------------------------------------------------------------- compaction

// Compaction works on 8 consecutive vertices (one byte of the bitmap) per
// thread: a workgroup covers kCompactNT words = 32 * kCompactNT vertices in
// four rounds of 8 * kCompactNT.  The in-degrees come from one read of the 9
// CSC offsets around the byte (coalesced across the wave).
constexpr int kCompactRounds = 4;

struct ByteWork {
  int v0;         // first vertex of the byte
  uint32_t bits;  // its 8 bits
  int cnt, deg;   // set bits and the sum of their in-degrees
  int o[9];       // off_in[v0 .. v0 + 8] (clamped to nV), valid when bits != 0
};

__device__ __forceinline__ ByteWork byte_work(const uint32_t* __restrict__ bitmap, int n_words,
                                              int nV, const int32_t* __restrict__ off_in,
                                              int need_deg, int byte_index) {
  ByteWork b;
  b.v0 = byte_index * 8;
  const int wi = byte_index >> 2;
  const uint32_t word = (wi < n_words) ? bitmap[wi] : 0u;
  b.bits = (word >> ((byte_index & 3) * 8)) & 0xFFu;
  b.cnt = __popc(b.bits);
  b.deg = 0;
  if (need_deg && b.bits) {
#pragma unroll
    for (int j = 0; j < 9; ++j) b.o[j] = off_in[min(b.v0 + j, nV)];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if ((b.bits >> j) & 1u) b.deg += b.o[j + 1] - b.o[j];  // bit set => v0 + j < nV
  }
  return b;
}

// One-byte degrees: deg[v] = min(degree, kDegSat) for v < nV, 0 for the
// padding up to n_pad.
[[maybe_unused]] static __global__ void k_degree8(int nV, int n_pad,
                                                  const int32_t* __restrict__ off_in,
                                                  const int32_t* __restrict__ off_out,
                                                  uint8_t* __restrict__ din,
                                                  uint8_t* __restrict__ dout) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n_pad; v += gridDim.x * blockDim.x) {
    int a = 0, b = 0;
    if (v < nV) {
      a = min(off_in[v + 1] - off_in[v], kDegSat);
      b = min(off_out[v + 1] - off_out[v], kDegSat);
    }
    din[v] = (uint8_t)a;
    dout[v] = (uint8_t)b;
  }
}

// nextIter's compaction in one launch.  Each block of kCompactNT words first
// folds the push pass's hub flags into its words (when the push ran: the same
// rule as the push kernels, on the same counters; the flags are cleared),
// counts the frontier's vertices and in-degrees, keeps its counts in
// blk[block], moves its words into the active bitmap `act_bits` (the next
// iteration's active set) and clears them in the frontier and in `chg_clear`
// (the changed bitmap the next apply fills), so that scatterActivate starts
// from an empty frontier.  Blocks form groups of kApplyGroup: each adds one
// packed word to its group (done count in bits 52 and up, vertices in bits
// 32-51, in-degrees in bits 0-31); the block that completes a group keeps
// the group's counts in fin.grp (for the list writer), adds them to fin.tot
// and then counts the group done, the second atomic issued only once the
// first has returned (its operand depends on the returned value), so that no
// fence is needed (a device-scope fence writes L2 back on this GPU); the
// block that completes the last group takes the totals out of fin.tot and
// publishes them: the
// active set's size and in-degrees, the changed-list counter's last value
// (reset for the next apply), the applies that returned true, written to the
// host's mapped copy and then `seq` to `host_seq`, so that the host can go on
// before the rest of the stream has finished.
struct HubFold {
  uint8_t* flags = nullptr;  // nullptr: no push pass in this iteration
  int n_hub = 0;
  int mode = 0;
  long long threshold = 0;
  long long fuse_max = -1;
};
struct CompactFinal {
  unsigned long long* gsum;  // [groups * kGroupStride], zero between launches
  int2* grp;                 // [groups]
  unsigned long long* tot;   // (vertices << 32 | in-degrees), zero between launches
  unsigned int* gdone;       // zero between launches
  DevCounters* ctr;
  int32_t* escan;            // nullptr: no in-degree scan end to write
  int mode;
  long long threshold;
  DevCounters* host;
  uint32_t* host_seq;
  uint32_t seq;
  unsigned long long* applied;
};
constexpr int kCompactDoneShift = 52;
static_assert((long long)kApplyGroup * 8 * kCompactRounds * kCompactNT <
                  (1ll << (kCompactDoneShift - 32)),
              "a group's vertex count fits below the done count");

[[maybe_unused]] static __global__ __launch_bounds__(kCompactNT) void k_compact_count(
    uint32_t* __restrict__ bitmap, int n_words, int nV, const int32_t* __restrict__ off_in,
    const uint8_t* __restrict__ deg8, int need_deg, int2* __restrict__ blk, uint32_t* __restrict__ act_bits,
    uint32_t* __restrict__ chg_clear, HubFold hub, CompactFinal fin) {
  __shared__ int s_a[kCompactNT / kWave], s_b[kCompactNT / kWave];
  __shared__ uint32_t s_word[kCompactNT];
  __shared__ int s_last;
  static_assert(kCompactRounds * kCompactNT / 4 == kCompactNT, "one word per thread");
  const int wi = blockIdx.x * kCompactNT + threadIdx.x;  // this thread's word
  uint32_t myword = wi < n_words ? bitmap[wi] : 0u;
  if (hub.flags) {  // uniform: every block reads the same counters
    const unsigned long long c = fin.ctr->chg;
    if (!(use_pull(c, hub.mode, hub.threshold) || fused_scatter(c, hub.fuse_max)) &&
        wi < hub.n_hub / 32) {
      uint4* t = reinterpret_cast<uint4*>(hub.flags + (size_t)wi * 32);
      const uint4 a = t[0], b = t[1];
      const uint32_t q[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
      uint32_t m = 0;
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) m |= ((q[i] >> (8 * j)) & 1u) << (4 * i + j);
      if (m) {
        myword |= m;
        t[0] = make_uint4(0, 0, 0, 0);
        t[1] = make_uint4(0, 0, 0, 0);
      }
    }
  }
  s_word[threadIdx.x] = myword;
  __syncthreads();
  int cnt = 0, deg = 0;
  uint32_t rbits[kCompactRounds];
#pragma unroll
  for (int r = 0; r < kCompactRounds; ++r) {
    const int bi = (blockIdx.x * kCompactRounds + r) * kCompactNT + threadIdx.x;
    const uint32_t word = s_word[r * (kCompactNT / 4) + threadIdx.x / 4];
    rbits[r] = (word >> ((bi & 3) * 8)) & 0xFFu;
    cnt += __popc(rbits[r]);
  }
  if (need_deg) {
    // each round's byte: its 8 vertices' one-byte in-degrees, one 8-byte
    // load (a set bit means bi < 4 * n_words, and deg8 covers every word's
    // vertices), all rounds' loads issued before the first is used; the
    // offsets only for a degree of kDegSat or more
    uint2 d[kCompactRounds];
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
      const int bi = (blockIdx.x * kCompactRounds + r) * kCompactNT + threadIdx.x;
      d[r] = rbits[r] ? *reinterpret_cast<const uint2*>(deg8 + (size_t)bi * 8) : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
      const int v0 = ((blockIdx.x * kCompactRounds + r) * kCompactNT + threadIdx.x) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if ((rbits[r] >> j) & 1u) {  // bit set => v0 + j < nV
          const int dj = (int)(((j < 4 ? d[r].x : d[r].y) >> (8 * (j & 3))) & 0xFFu);
          deg += dj == kDegSat ? off_in[v0 + j + 1] - off_in[v0 + j] : dj;
        }
      }
    }
  }
  cnt = wave_reduce_add(cnt);
  deg = wave_reduce_add(deg);
  if (lane_id() == 0) {
    s_a[threadIdx.x / kWave] = cnt;
    s_b[threadIdx.x / kWave] = deg;
  }
  if (wi < n_words) {
    act_bits[wi] = myword;
    bitmap[wi] = 0u;
    if (chg_clear) chg_clear[wi] = 0u;
  }
  __syncthreads();
  const int nb = (int)gridDim.x;
  const int n_groups = (nb + kApplyGroup - 1) / kApplyGroup;
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int i = 0; i < kCompactNT / kWave; ++i) {
      a += s_a[i];
      b += s_b[i];
    }
    blk[blockIdx.x] = make_int2(a, b);
    const int grp = (int)blockIdx.x / kApplyGroup;
    const unsigned members = (unsigned)min(kApplyGroup, nb - grp * kApplyGroup);
    const unsigned long long mine = (1ull << kCompactDoneShift) |
                                    ((unsigned long long)(uint32_t)a << 32) | (uint32_t)b;
    const unsigned long long old = atomicAdd(&fin.gsum[grp * kGroupStride], mine);
    int last = 0;
    if ((unsigned)(old >> kCompactDoneShift) == members - 1) {
      const unsigned long long g = (old + mine) & ((1ull << kCompactDoneShift) - 1);
      fin.gsum[grp * kGroupStride] = 0ull;  // every block of the group has added
      fin.grp[grp] = make_int2((int)(g >> 32), (int)(uint32_t)g);
      // (vertices < 2^31: bit 63 of the returned total is 0, but the done
      // count's add depends on it and waits for the first add to return)
      const unsigned long long t = atomicAdd(fin.tot, g);
      last = atomicAdd(fin.gdone, 1u + (unsigned)(t >> 63)) == (unsigned)(n_groups - 1) ? 1 : 0;
    }
    s_last = last;
  }
  __syncthreads();
  if (!s_last) return;  // uniform
  unsigned long long my_applied = 0;
  if (threadIdx.x < kAppliedSlots) {
    my_applied = fin.applied[threadIdx.x * kAppliedStride];
    fin.applied[threadIdx.x * kAppliedStride] = 0;
  }
  my_applied = wave_reduce_add(my_applied);  // slots 0..63: wave 0
  if (threadIdx.x == 0) {
    // every group's add was performed before its done count
    const unsigned long long t = atomicExch(fin.tot, 0ull);
    const long long ta = (long long)(t >> 32), tb = (long long)(uint32_t)t;
    DevCounters* ctr = fin.ctr;
    *fin.gdone = 0u;
    ctr->n_active = (int32_t)ta;
    ctr->w_edges = (int32_t)tb;
    if (fin.escan) fin.escan[ta] = (int32_t)tb;
    ctr->chg_last = ctr->chg;
    ctr->pull = use_pull(ctr->chg, fin.mode, fin.threshold) && !ctr->fused ? 1 : 0;
    ctr->chg = 0;
    ctr->applied_last = my_applied;
    *fin.host = *ctr;
    __threadfence_system();
    __hip_atomic_store(fin.host_seq, fin.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The exclusive (vertices, in-degrees) before compaction block b: the groups
// before its group and the blocks of its group before it.  Wave 0 computes
// it into *out; every thread of the block calls this.
__device__ __forceinline__ void compact_block_base(const int2* __restrict__ blk,
                                                   const int2* __restrict__ grp, int b,
                                                   int2* out) {
  if (threadIdx.x < kWave) {
    const int g = b / kApplyGroup;
    long long a = 0, d = 0;
    for (int i = threadIdx.x; i < g; i += kWave) {
      a += grp[i].x;
      d += grp[i].y;
    }
    for (int i = g * kApplyGroup + (int)threadIdx.x; i < b; i += kWave) {
      a += blk[i].x;
      d += blk[i].y;
    }
    a = wave_reduce_add(a);
    d = wave_reduce_add(d);
    if (threadIdx.x == 0) *out = make_int2((int)a, (int)d);
  }
  __syncthreads();
}

// Writes the ascending active list of the active bitmap (as k_compact_count
// left it), its exclusive in-degree scan and the CSC base per row.  Run when a
// phase needs the list (the merge-path gather, the streamed gather, a flush
// of pending candidates, apply after such a gather).
[[maybe_unused]] static __global__ __launch_bounds__(kCompactNT) void k_compact_write(
    const uint32_t* __restrict__ bitmap, int n_words, int nV, const int32_t* __restrict__ off_in,
    int need_deg, const int2* __restrict__ blk, const int2* __restrict__ grp,
    int32_t* __restrict__ active, int32_t* __restrict__ escan, int32_t* __restrict__ cbase) {
  // a round's entries (at most 8 per thread) are staged here and written out
  // coalesced: each thread's own entries are consecutive, a wave's are not
  constexpr int kRoundMax = 8 * kCompactNT;
  __shared__ unsigned long long s_scan[kCompactNT / kWave];
  __shared__ int32_t s_act[kRoundMax], s_esc[kRoundMax], s_cb[kRoundMax];
  // a block without active bits has nothing to write
  if (blk[blockIdx.x].x == 0) return;  // uniform
  __shared__ int2 s_base;
  compact_block_base(blk, grp, (int)blockIdx.x, &s_base);
  const int2 base = s_base;
  int pos_run = base.x, deg_run = base.y;
  for (int r = 0; r < kCompactRounds; ++r) {
    const int bi = (blockIdx.x * kCompactRounds + r) * kCompactNT + threadIdx.x;
    const ByteWork b = byte_work(bitmap, n_words, nV, off_in, need_deg, bi);
    // (deg << 16 | cnt): cnt <= 8 per thread and <= 2048 per round
    unsigned long long tot;
    const unsigned long long ex = block_excl_scan_add<kCompactNT>(
        ((unsigned long long)(uint32_t)b.deg << 16) | (unsigned long long)b.cnt, s_scan, &tot);
    const int round_start = pos_run, round_cnt = (int)(tot & 0xFFFFu);
    int pos = (int)(ex & 0xFFFFu);  // within the round
    int dpos = deg_run + (int)(ex >> 16);
    pos_run += round_cnt;
    deg_run += (int)(tot >> 16);
    if (round_cnt == 0) continue;  // block-uniform
    uint32_t bits = b.bits;
    while (bits) {
      const int j = __builtin_ctz(bits);
      bits &= bits - 1;
      s_act[pos] = b.v0 + j;
      if (need_deg) {
        s_esc[pos] = dpos;
        s_cb[pos] = b.o[j] - dpos;
        dpos += b.o[j + 1] - b.o[j];
      }
      ++pos;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < round_cnt; i += kCompactNT) {
      active[round_start + i] = s_act[i];
      if (need_deg) {
        escan[round_start + i] = s_esc[i];
        cbase[round_start + i] = s_cb[i];
      }
    }
    __syncthreads();  // the next round stages into the same arrays
  }
}

[[maybe_unused]] static __global__ void k_set_active_range(int start, int n, const int32_t* __restrict__ off_in,
                                   int32_t* __restrict__ active, int32_t* __restrict__ escan,
                                   int32_t* __restrict__ cbase,
                                   DevCounters* __restrict__ ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int base = off_in[start];
  if (i < n) {
    active[i] = start + i;
    cbase[i] = base;
  }
  escan[i] = off_in[start + i] - base;
  if (i == n) {
    ctr->n_active = n;
    ctr->w_edges = off_in[start + n] - base;
    ctr->chg = 0;
  }
}

[[maybe_unused]] static __global__ void k_set_counters(DevCounters* __restrict__ ctr, int n, int w) {
  ctr->n_active = n;
  ctr->w_edges = w;
  ctr->chg = 0;
}

// Gathers n elements of `words` 4-byte words: out[i] = in[eid[i]] (eid may be
// nullptr: identity; an eid of -1 writes zeros).  Moves edge data from
// edge-list order into CSC order and into the blocked gather's slot order,
// and vertex data between user order and the degree-ordered layout.
[[maybe_unused]] static __global__ void k_permute_words(const uint32_t* __restrict__ in, const int32_t* __restrict__ eid,
                                int64_t n, int words, uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = eid ? (int64_t)eid[i] : i;
  for (int k = 0; k < words; ++k) out[i * words + k] = s >= 0 ? in[s * words + k] : 0u;
}

// ------------------------------------------------ out-of-core shards
//
// With the edge arrays in host memory (gr_engine_set_streaming), vertex-sized
// state stays on the device and every iteration streams the shards it needs:
// the CSC slices of the destination ranges that hold active rows (gather) and
// the CSR slices of the source ranges that hold changed vertices (activation).

// pos[s] = first active-list index with vertex >= bound[s]; wsum[s] = escan
// there (the active list is ascending).
[[maybe_unused]] static __global__ void k_shard_positions(int n_bounds, const int32_t* __restrict__ bound,
                                  const int32_t* __restrict__ active, int n_active,
                                  const int32_t* __restrict__ escan, int32_t* __restrict__ pos,
                                  int32_t* __restrict__ wsum) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_bounds) return;
  const int b = bound[s];
  int lo = 0, hi = n_active;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (active[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  pos[s] = lo;
  wsum[s] = escan[lo];
}

// The active sub-list [a0, a0 + n] of one shard as a list of its own: its
// in-degree scan from 0 and CSC bases relative to the shard's first edge e0.
[[maybe_unused]] static __global__ void k_shard_rebase(int n, const int32_t* __restrict__ escan,
                               const int32_t* __restrict__ cbase, int a0, int e0,
                               int32_t* __restrict__ escan_sub, int32_t* __restrict__ cbase_sub) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int base = escan[a0];
  escan_sub[i] = escan[a0 + i] - base;
  if (i < n) cbase_sub[i] = cbase[a0 + i] + base - e0;
}

// cnt[s] = changed vertices (bits set) in [bound[s], bound[s + 1]); one
// workgroup per shard.
[[maybe_unused]] static __global__ void k_shard_changed(int n_shards, const int32_t* __restrict__ bound,
                                const uint32_t* __restrict__ bits, int32_t* __restrict__ cnt) {
  __shared__ int s_sum;
  const int s = blockIdx.x;
  if (s >= n_shards) return;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  const int v0 = bound[s], v1 = bound[s + 1];
  int c = 0;
  for (int v = v0 + threadIdx.x; v < v1; v += blockDim.x) c += (bits[v >> 5] >> (v & 31)) & 1u;
  atomicAdd(&s_sum, c);
  __syncthreads();
  if (threadIdx.x == 0) cnt[s] = s_sum;
}

// Activation from one staged CSR shard: every out-neighbour of a changed
// vertex in [u0, u1) gets its frontier bit; dsts_stage holds the shard's
// targets from CSR position e0 on.  One wave per 64 sources.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_shard_push(int u0, int u1, int e0,
                                                    const int32_t* __restrict__ off_out,
                                                    const int32_t* __restrict__ dsts_stage,
                                                    const uint32_t* __restrict__ chg_bits,
                                                    uint32_t* __restrict__ bitmap) {
  const int lane = lane_id();
  const int nw = gridDim.x * blockDim.x / kWave;
  for (int base = u0 + (blockIdx.x * blockDim.x + threadIdx.x) / kWave * kWave; base < u1;
       base += nw * kWave) {
    const int u = base + lane;
    const bool ch = u < u1 && ((chg_bits[u >> 5] >> (u & 31)) & 1u);
    uint64_t todo = __ballot(ch);
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int b = off_out[base + j] - e0, e = off_out[base + j + 1] - e0;
      for (int p = b + lane; p < e; p += kWave) {
        const int v = dsts_stage[p];
        const uint32_t bit = 1u << (v & 31);
        if (!(bitmap[v >> 5] & bit)) atomicOr(&bitmap[v >> 5], bit);
      }
    }
  }
}

// out[eid[i]] = in[i] (the inverse of k_permute_words), n elements of `words` words.
[[maybe_unused]] static __global__ void k_unpermute_words(const uint32_t* __restrict__ in, const int32_t* __restrict__ eid,
                                  int64_t n, int words, uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t d = eid[i];
  for (int k = 0; k < words; ++k) out[d * words + k] = in[i * words + k];
}

// ================================================================ host side

struct EngineBase {
  virtual ~EngineBase() {}
  virtual int bind(gr_graph* g, bool owns, void* vdata, size_t velem, const void* edata,
                   size_t eelem, bool edata_on_device, bool stream_mode = false) = 0;
  // setGraph with the edge arrays kept in host memory (stream_max_edges > 0)
  virtual int bind_streaming(int32_t nV, void* vdata, size_t velem, int32_t nE, const void* edata,
                             size_t eelem, const int32_t* srcs, const int32_t* dsts) = 0;
  virtual int set_active(int32_t start, int32_t end) = 0;
  virtual int gather() = 0;
  virtual int apply() = 0;
  virtual int scatter_activate(bool have_scatter) = 0;
  virtual int next_iter() = 0;
  virtual int get_results() = 0;
  virtual int read_vertex_data(void* out, size_t velem) = 0;
  virtual int read_edge_data(void* out, size_t eelem) = 0;
  virtual int reset_vertex_data() = 0;
  virtual int set_delta_gather(int mode) = 0;
  virtual int sync_stats(gr_engine_stats* out) = 0;
  virtual size_t vertex_size() const = 0;
  virtual size_t edge_size() const = 0;  // 0 for an empty EdgeData
  // after gr_engine_run's loop: drops an apply its last nextIter queued (it
  // ran over the empty active set the loop ended on)
  virtual int end_run() { return GR_OK; }

  int32_t n_active = 0;
  int32_t w_edges = 0;
  int32_t iteration = 0;
  bool bound = false;
  // set by gr_engine_run for the length of its loop: no caller sees the
  // state between nextIter and the next apply, so nextIter may queue that
  // apply before the host has the next frontier's size
  bool run_mode = false;
  int dir_mode = 0;
  double dir_alpha = 15.0;
  int gather_vt = 8;  // merge-path items per thread in the gather tiles (4, 8 or 16)
  // 1: the engine works on the degree-ordered copy of the graph; 0: on the
  // graph as given; 2: the copy when the degrees are skewed (graph_skewed)
  int vertex_order = 2;
  int64_t stream_max_edges = 0;  // > 0: out-of-core shards of at most this many edges
  double fuse_share = 0.4;        // fused activation + push gather up to this share of E
  // blocked gather (programs that qualify) for gathers over at least this
  // share of E; <= 0: never (gr_engine_set_blocked_gather)
  double blocked_share = 0.5;
  // adaptive delta gather: push the changed out-edges below this share of the
  // gather's in-edges (gr_engine_set_push_share)
  double push_share = 0.6;
  bool timing = false;
  gr_engine_stats stats{};
};

template <class P>
struct Engine final : EngineBase {
  typedef typename P::VertexData VD;
  typedef typename P::EdgeData ED;
  typedef typename P::GatherResult G;
  static constexpr bool kEdgeEmpty = std::is_empty<ED>::value;
  // Programs with a scatter keep ONE copy of their edge data, in edge-list
  // order (the scatter changes it; the gather reads it through csc_eid).
  static constexpr bool kScatter = P::kHasScatter && !std::is_empty<typename P::EdgeData>::value;

  hipStream_t stream = nullptr;
  gr_graph* g = nullptr;       // the layout the kernels run on (user_g or its ordered copy)
  gr_graph* user_g = nullptr;  // the graph as given
  const int32_t* d_perm = nullptr;  // layout id -> user id (nullptr: same ids)
  const int32_t* d_inv = nullptr;   // user id -> layout id
  bool owns_graph = false;
  void* host_vdata = nullptr;
  VD* d_vd = nullptr;
  VD* d_vd_init = nullptr;  // vertex data as given at bind time
  VD* d_vd_user = nullptr;  // staging in user order (ordered layout only)
  ED* d_ed = nullptr;  // CSC order (edge-list order for programs with a scatter)
  ED* d_ed_csr = nullptr;  // CSR order (delta-gather programs: push relax)
  G* d_cand = nullptr;     // push-relax candidates per vertex (gatherZero between gathers)
  // Blocked gather: qualifying programs, the graph's source-listed row blocks,
  // the edge data in their order, and the vertex-indexed gather values.
  static constexpr bool kBlockedMin = std::is_same<G, int>::value && delta_gather_v<P> &&
                                      !P::kGatherUsesDst && !P::kHasScatter;
  static constexpr bool kBlockedSum = std::is_same<G, float>::value && fixed_point_sum_v<P> &&
                                      !P::kGatherUsesDst && !P::kHasScatter;
  static constexpr bool kBlocked = kBlockedMin || kBlockedSum;
  // in- and out-degrees of the vertices in one byte each, kDegSat standing
  // for kDegSat or more (then the offsets are read)
  uint8_t* d_deg_in8 = nullptr;
  uint8_t* d_deg_out8 = nullptr;
  const gr_blocked* blk = nullptr;
  int blk_rows = kBlockRows;  // the blocks' row limit (kBlockRows or kBlockRowsWide)
  ED* d_ed_blk = nullptr;
  G* d_gdense = nullptr;   // rows from blk->rows_end on stay gatherZero
  bool dense_pending = false;  // apply reads d_gdense[v] (the last gather was blocked)
  // Fused activation: scatterActivate runs the next blocked gather over every
  // row, which also marks the rows with a changed in-neighbour (the pull);
  // the next gather only collects its values.  Fixed-point sum programs do
  // so after a blocked gather; min programs whenever the pull would run (the
  // kernel decides on the device and records in ctr->fab whether it ran).
  bool last_gather_blocked = false;
  bool fab_pending = false;
  bool fab_launched = false;  // scatterActivate enqueued the min programs' fused pass
  source_value_t<P>* d_srcval = nullptr;  // P::sourceValue of every vertex (programs with one)
  // ---- out-of-core mode: edge arrays in pinned host memory, streamed per shard
  bool streaming = false;
  int32_t* h_srcs = nullptr;  // CSC sources
  int32_t* h_dsts = nullptr;  // CSR targets
  ED* h_ed = nullptr;         // edge data in CSC order
  std::vector<int32_t> h_csc_eid;         // CSC position -> edge-list index
  std::vector<int32_t> csc_vb, csr_vb;    // shard vertex bounds
  int32_t* d_csc_vb = nullptr;
  int32_t* d_csr_vb = nullptr;
  int32_t* d_spos = nullptr;   // per CSC shard: first active-list index, escan there
  int32_t* d_swsum = nullptr;
  int32_t* d_scnt = nullptr;   // per CSR shard: changed vertices
  int32_t* h_sbuf = nullptr;   // pinned host copy of those
  int32_t* d_stage_src[2] = {nullptr, nullptr};
  ED* d_stage_ed[2] = {nullptr, nullptr};
  int32_t* d_stage_dst[2] = {nullptr, nullptr};
  int32_t* d_escan_sub = nullptr;
  int32_t* d_cbase_sub = nullptr;
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_copied[2] = {nullptr, nullptr};
  hipEvent_t ev_used[2] = {nullptr, nullptr};
  int32_t* d_active = nullptr;
  int32_t* d_escan = nullptr;
  int32_t* d_cbase = nullptr;  // per active row: CSC offset minus its escan value
  G* d_gtmp = nullptr;
  int32_t* d_chg_v = nullptr;
  int32_t* d_chg_e = nullptr;
  uint32_t* d_bitmap = nullptr;   // next frontier (scatterActivate marks it; empty otherwise)
  // The active set after nextIter is the bitmap d_act_bits; the ascending list
  // (d_active, d_escan, d_cbase) is written from it only when a phase needs it
  // (ensure_list); setActive writes the list itself.
  uint32_t* d_act_bits = nullptr;
  bool list_valid = true;
  // Deferred changed list (large bitmap applies): per-chunk counts, scanned
  // in place; the list itself is written when a phase needs it.
  unsigned long long* d_cc = nullptr;
  unsigned long long* d_sc = nullptr;
  unsigned long long* d_sc_final = nullptr;
  int n_apply_chunks = 0;
  bool chg_pending = false;
  bool chg_cond = false;  // written only if the blocked gather did not run
  bool pre_applied = false;  // nextIter queued the next iteration's apply
  bool gated_apply = false;  // ... gated on the device by ctr->fab (kBlockedMin)
  // source_flag_v programs: the sign bits of d_srcval hold at most the last
  // flag-keeping apply's changed set (src_flags_ok), exactly it (gather_flags:
  // the fused blocked sum may read them instead of the changed bitmap)
  bool src_flags_ok = true;
  bool gather_flags = false;
  uint32_t* d_chg_bits = nullptr;  // vertices changed by this iteration's apply
  uint32_t* d_chg_prev = nullptr;  // ... by the previous iteration's apply
  int2* d_blk = nullptr;
  // nextIter's compaction: per-group sums (zero between launches), group
  // counts, groups done (zero between launches)
  unsigned long long* d_cgsum = nullptr;
  int2* d_cgrp = nullptr;
  unsigned long long* d_cgtot = nullptr;
  unsigned int* d_cgdone = nullptr;
  HubFold hub;  // hub marks a push left for the next compaction to fold
  int32_t* d_part = nullptr;
  int32_t* d_push_part = nullptr;
  Part<G>* d_carry = nullptr;
  Part<G>* d_head = nullptr;
  DevCounters* d_ctr = nullptr;
  unsigned long long* d_applied = nullptr;  // kAppliedSlots spread counters
  DevCounters* h_ctr = nullptr;  // pinned, mapped, coherent
  DevCounters* d_hctr = nullptr; // h_ctr as the device addresses it
  uint32_t* h_seq = nullptr;     // last counters' sequence number (pinned, mapped, coherent)
  uint32_t* d_hseq = nullptr;
  uint32_t ctr_seq = 0;
  int n_words = 0, n_cblocks = 0, max_tiles = 0, max_push_tiles = 0;
  int32_t* d_heavy = nullptr;  // vertices with in-degree > kPullHeavy (pull mode)
  uint8_t* d_hub_flags = nullptr;  // push marks of vertices below n_hub (zero between scatters)
  // Out-degree histogram by log2 bucket (vertices, out-edges): bounds the
  // out-edges of any k vertices, so that scatterActivate does not launch the
  // kernels whose device-side condition the changed vertices cannot meet.
  int64_t deg_cnt[32] = {}, deg_sum[32] = {};
  int n_applies = 0;  // apply calls since the active set was last built
  int n_hub = 0;                   // multiple of 32, at most kHubVertices
  int n_heavy = 0;
  // Delta gather (programs with kDeltaGather): valid once an all-active
  // iteration has run and every later active set came from scatterActivate.
  static constexpr bool kDelta = delta_gather_v<P>;
  // 0 off; 1 adaptive (push below push_share, else full); 2 delta
  // pull whenever valid; 3 push whenever valid
  int delta_mode = 1;
  int64_t chg_count = 0, chg_edges = 0;  // changed list of the previous iteration
  bool delta_chain = false;  // the invariant holds for the current state
  bool delta_ready = false;  // d_chg_prev holds the previous iteration's changes
  // Share of the next gather's in-edges whose source changed (E_out(C) / E_in(A)).
  // Measured per iteration on RMAT-24 SSSP: pushing the changed out-edges is
  // the fastest gather below a share of about 0.6 (1.6 ms vs 1.7 ms pull-delta
  // and 1.8 ms full at 0.56; 0.7 ms vs 1.3 and 1.7 at 0.16), the full gather
  // above it (1.8 ms vs 1.85 push and 2.1 pull-delta at 0.65; 1.8 vs 2.4 and
  // 2.5 at 0.94); the pull-delta gather is never the fastest there.
  double changed_share = 1.0;
  // Fused activation + push gather when the changed vertices' out-edges are at
  // most fuse_share of E (adaptive mode; gr_engine_set_fuse_share): below it,
  // the separate activation (pull scans of whole rows, or a push) plus the
  // next push relax cost more than one relax pass that also marks the targets.
  bool fuse_launched = false;  // scatterActivate enqueued the fused pass
  bool fused_pending = false;  // d_cand holds the next gather's values
  // A push gather leaves the active rows' values in d_cand: apply reads and
  // resets them; any other phase first resets them (collect_flush).
  bool collect_pending = false;
  // Inert sources (programs with kInertSources): with every bound edge passing
  // P::edgeKeepsInert, a gather without a previous changed set reads only the
  // non-inert sources (SSSP's first all-active gather: the start vertex).
  static constexpr bool kInert = kDelta && inert_sources_v<P> && !kEdgeEmpty;
  bool inert_ok = false;
  bool it_gather = false, it_apply = false, it_scatter = false;
  // hipEvent start/stop pairs per timed phase
  enum { kPhGather = 0, kPhScatter, kPhApply, kPhNext, kPhPush, kPhInert, kPhases };
  std::vector<hipEvent_t> ev[kPhases];
  hipEvent_t ev_open = nullptr;

  ~Engine() override { release(); }

  size_t vertex_size() const override { return sizeof(VD); }
  size_t edge_size() const override { return kEdgeEmpty ? 0 : sizeof(ED); }

  void release() {
    // nextIter returns while the compaction scan may still run: drain the
    // stream before its buffers go
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : {(void*)d_ed_blk, (void*)d_gdense})
      if (p) (void)hipFree(p);
    d_ed_blk = nullptr;
    d_gdense = nullptr;
    blk = nullptr;
    dense_pending = false;
    void* ptrs[] = {d_hub_flags, d_srcval, d_cbase,    d_heavy,  d_vd,     d_vd_init,
                    d_vd_user,   d_ed,     d_ed_csr,   d_cand,   d_active, d_escan,
                    d_gtmp,      d_chg_v,  d_chg_e,    d_bitmap, d_chg_bits, d_chg_prev,
                    d_blk,       d_part,   d_push_part, d_carry, d_head,   d_ctr,
                    d_applied,   d_act_bits, d_cc,     d_sc,     d_sc_final, d_cgsum,
                    d_cgrp,      d_cgtot,  d_cgdone, d_deg_in8, d_deg_out8};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (h_ctr) (void)hipHostFree(h_ctr);
    if (h_seq) (void)hipHostFree(h_seq);
    {
      void* sptrs[] = {d_csc_vb,       d_csr_vb,       d_spos,         d_swsum,
                       d_scnt,         d_stage_src[0], d_stage_src[1], d_stage_ed[0],
                       d_stage_ed[1],  d_stage_dst[0], d_stage_dst[1], d_escan_sub,
                       d_cbase_sub};
      for (void* p : sptrs)
        if (p) (void)hipFree(p);
      void* hptrs[] = {h_srcs, h_dsts, h_ed, h_sbuf};
      for (void* p : hptrs)
        if (p) (void)hipHostFree(p);
      for (int b = 0; b < 2; ++b) {
        if (ev_copied[b]) (void)hipEventDestroy(ev_copied[b]);
        if (ev_used[b]) (void)hipEventDestroy(ev_used[b]);
        ev_copied[b] = ev_used[b] = nullptr;
        d_stage_src[b] = d_stage_dst[b] = nullptr;
        d_stage_ed[b] = nullptr;
      }
      if (copy_stream) (void)hipStreamDestroy(copy_stream);
      copy_stream = nullptr;
      d_csc_vb = d_csr_vb = d_spos = d_swsum = d_scnt = d_escan_sub = d_cbase_sub = nullptr;
      h_srcs = h_dsts = h_sbuf = nullptr;
      h_ed = nullptr;
      h_csc_eid.clear();
      csc_vb.clear();
      csr_vb.clear();
      streaming = false;
    }
    for (auto& v : ev) {
      for (hipEvent_t e : v) (void)hipEventDestroy(e);
      v.clear();
    }
    if (stream) (void)hipStreamDestroy(stream);
    if (owns_graph && user_g) graph_free(user_g);
    user_g = nullptr;
    d_perm = d_inv = nullptr;
    d_vd = nullptr;
    d_vd_init = nullptr;
    d_vd_user = nullptr;
    d_heavy = nullptr;
    d_hub_flags = nullptr;
    n_hub = 0;
    n_heavy = 0;
    d_ed = nullptr;
    d_ed_csr = nullptr;
    d_cand = nullptr;
    d_srcval = nullptr;
    d_cbase = nullptr;
    d_active = d_escan = d_chg_v = d_chg_e = d_part = d_push_part = nullptr;
    d_gtmp = nullptr;
    d_bitmap = d_chg_bits = d_chg_prev = d_act_bits = nullptr;
    d_cc = d_sc = d_sc_final = d_cgsum = d_cgtot = nullptr;
    d_cgrp = nullptr;
    d_cgdone = nullptr;
    d_deg_in8 = d_deg_out8 = nullptr;
    hub = HubFold();
    chg_pending = chg_cond = false;
    list_valid = true;
    d_blk = nullptr;
    d_carry = d_head = nullptr;
    d_ctr = h_ctr = d_hctr = nullptr;
    d_applied = nullptr;
    h_seq = d_hseq = nullptr;
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
           size_t eelem, bool edata_on_device, bool stream_mode = false) override {
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
    user_g = graph;
    g = graph;
    owns_graph = owns;
    host_vdata = vdata;
    streaming = stream_mode;  // edge arrays stay in host memory (bind_streaming)
    if (graph->nV > 0 && !streaming &&
        (vertex_order == 1 || (vertex_order == 2 && graph_skewed(graph)))) {
      GR_TRY(graph_ordered(graph, &g));
      d_perm = graph->perm;
      d_inv = graph->inv;
    }
    const int64_t V = g->nV, E = g->nE;
    if (V + E >= ((int64_t)1 << 31) - kNT * 16)
      return set_error(GR_ERR_INVALID, "graph too large for int32 merge-path coordinates");
    GR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    n_words = (int)((V + 31) / 32);
    n_cblocks = (n_words + kCompactNT - 1) / kCompactNT;
    // sized for the smallest tile (gather_vt = 4)
    max_tiles = (int)((V + E + kNT * 4 - 1) / (kNT * 4)) + 1;
    max_push_tiles = (int)((E + kTile - 1) / kTile) + 1;
    GR_TRY(dalloc(&d_vd, V));
    GR_TRY(dalloc(&d_vd_init, V));
    GR_TRY(dalloc(&d_active, V));
    GR_TRY(dalloc(&d_escan, V + 1));
    GR_TRY(dalloc(&d_cbase, V + 1));
    GR_TRY(dalloc(&d_gtmp, V));
    GR_TRY(dalloc(&d_chg_v, V));
    GR_TRY(dalloc(&d_chg_e, V));
    GR_TRY(dalloc(&d_bitmap, n_words + 1));
    GR_TRY(dalloc(&d_act_bits, n_words + 1));
    GR_TRY(dalloc(&d_chg_bits, n_words + 1));
    // one-byte degrees for the counting passes: every vertex of the bitmaps'
    // words has a byte (0 past nV)
    GR_TRY(dalloc(&d_deg_in8, (int64_t)n_words * 32));
    GR_TRY(dalloc(&d_deg_out8, (int64_t)n_words * 32));
    if (n_words > 0) {
      k_degree8<<<grid_1d((int64_t)n_words * 32, 256), 256, 0, stream>>>(
          g->nV, n_words * 32, g->off_in, g->off_out, d_deg_in8, d_deg_out8);
      GR_HIP(hipGetLastError());
    }
    GR_TRY(dalloc(&d_chg_prev, n_words + 1));
    n_apply_chunks = (int)(((int64_t)V + kApplyChunk - 1) / kApplyChunk);
    GR_TRY(dalloc(&d_cc, (int64_t)n_apply_chunks));
    {
      const int64_t groups = (n_apply_chunks + kApplyGroup - 1) / kApplyGroup;
      GR_TRY(dalloc(&d_sc, groups * kGroupStride));
      GR_TRY(dalloc(&d_sc_final, groups));
      GR_HIP(hipMemsetAsync(d_sc, 0, sizeof(unsigned long long) * groups * kGroupStride, stream));
    }
    GR_TRY(dalloc(&d_blk, n_cblocks));
    {
      const int64_t groups = (n_cblocks + kApplyGroup - 1) / kApplyGroup;
      GR_TRY(dalloc(&d_cgsum, groups * kGroupStride));
      GR_TRY(dalloc(&d_cgrp, groups));
      GR_TRY(dalloc(&d_cgtot, 1));
      GR_TRY(dalloc(&d_cgdone, 1));
      GR_HIP(hipMemsetAsync(d_cgtot, 0, sizeof(unsigned long long), stream));
      GR_HIP(hipMemsetAsync(d_cgsum, 0, sizeof(unsigned long long) * groups * kGroupStride,
                            stream));
      GR_HIP(hipMemsetAsync(d_cgdone, 0, sizeof(unsigned int), stream));
    }
    GR_TRY(dalloc(&d_part, max_tiles + 2));
    GR_TRY(dalloc(&d_push_part, max_push_tiles + 2));
    GR_TRY(dalloc(&d_carry, max_tiles + 1));
    GR_TRY(dalloc(&d_head, max_tiles + 1));
    GR_TRY(dalloc(&d_ctr, 1));
    GR_TRY(dalloc(&d_applied, kAppliedSlots * kAppliedStride));
    GR_HIP(hipHostMalloc((void**)&h_ctr, sizeof(DevCounters),
                         hipHostMallocMapped | hipHostMallocCoherent));
    GR_HIP(hipHostGetDevicePointer((void**)&d_hctr, h_ctr, 0));
    GR_HIP(hipHostMalloc((void**)&h_seq, sizeof(uint32_t),
                         hipHostMallocMapped | hipHostMallocCoherent));
    GR_HIP(hipHostGetDevicePointer((void**)&d_hseq, h_seq, 0));
    __atomic_store_n(h_seq, ctr_seq, __ATOMIC_RELEASE);
    GR_HIP(hipMemsetAsync(d_bitmap, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_act_bits, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_bits, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_prev, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_ctr, 0, sizeof(DevCounters), stream));
    list_valid = true;
    GR_HIP(hipMemsetAsync(d_applied, 0,
                          sizeof(unsigned long long) * kAppliedSlots * kAppliedStride, stream));
    {
      std::vector<int32_t> hv;
      for (int32_t v = 0; v < g->nV; ++v)
        if (g->h_off_in[v + 1] - g->h_off_in[v] > kPullHeavy) hv.push_back(v);
      n_heavy = (int)hv.size();
      GR_TRY(dalloc(&d_heavy, n_heavy));
      for (int b = 0; b < 32; ++b) deg_cnt[b] = deg_sum[b] = 0;
      for (int32_t v = 0; v < g->nV; ++v) {
        const int32_t d = g->h_off_out[v + 1] - g->h_off_out[v];
        if (d <= 0) continue;
        const int b = 31 - __builtin_clz((unsigned)d);
        deg_cnt[b] += 1;
        deg_sum[b] += d;
      }
      n_hub = (int)std::min<int64_t>(V / 32 * 32, kHubVertices);
      if (n_hub > 0) {
        GR_TRY(dalloc(&d_hub_flags, n_hub));
        GR_HIP(hipMemsetAsync(d_hub_flags, 0, (size_t)n_hub, stream));
      }
      if (n_heavy)
        GR_HIP(hipMemcpy(d_heavy, hv.data(), sizeof(int32_t) * n_heavy, hipMemcpyHostToDevice));
    }
    if (d_perm) {
      GR_TRY(dalloc(&d_vd_user, V));
      GR_HIP(hipMemcpyAsync(d_vd_user, vdata, sizeof(VD) * V, hipMemcpyHostToDevice, stream));
      to_layout(d_vd_user, d_vd);
    } else {
      GR_HIP(hipMemcpyAsync(d_vd, vdata, sizeof(VD) * V, hipMemcpyHostToDevice, stream));
    }
    GR_HIP(hipMemcpyAsync(d_vd_init, d_vd, sizeof(VD) * V, hipMemcpyDeviceToDevice, stream));
    if constexpr (source_value_v<P>) {
      GR_TRY(dalloc(&d_srcval, V));
      GR_TRY(refresh_source_values());
    }
    if constexpr (!kEdgeEmpty) {
      GR_TRY(dalloc(&d_ed, streaming ? 0 : E));
      if (E > 0 && !streaming) {
        const ED* src = static_cast<const ED*>(edata);
        ED* tmp = nullptr;
        if (!edata_on_device) {
          GR_TRY(dalloc(&tmp, E));
          GR_HIP(hipMemcpyAsync(tmp, edata, sizeof(ED) * E, hipMemcpyHostToDevice, stream));
          src = tmp;
        }
        if constexpr (kScatter) {
          GR_HIP(hipMemcpyAsync(d_ed, src, sizeof(ED) * E, hipMemcpyDeviceToDevice, stream));
        } else {
          k_permute_words<<<grid_1d(E, 256), 256, 0, stream>>>(
              reinterpret_cast<const uint32_t*>(src), g->csc_eid, E, (int)(sizeof(ED) / 4),
              reinterpret_cast<uint32_t*>(d_ed));
          GR_HIP(hipGetLastError());
        }
        if constexpr (kDelta) {
          GR_TRY(dalloc(&d_ed_csr, E));
          k_permute_words<<<grid_1d(E, 256), 256, 0, stream>>>(
              reinterpret_cast<const uint32_t*>(src), g->csr_eid, E, (int)(sizeof(ED) / 4),
              reinterpret_cast<uint32_t*>(d_ed_csr));
          GR_HIP(hipGetLastError());
        }
        GR_HIP(hipStreamSynchronize(stream));
        if (tmp) (void)hipFree(tmp);
      }
    }
    inert_ok = false;
    if constexpr (kInert) {
      if (!streaming) {  // the weights are checked on the device copy
        unsigned int* d_bad = nullptr;
        unsigned int h_bad = 1;
        GR_TRY(dalloc(&d_bad, 1));
        GR_HIP(hipMemsetAsync(d_bad, 0, sizeof(unsigned int), stream));
        if (E > 0)
          k_check_inert_edges<P><<<grid_1d(E, 256 * 16) < 8192 ? grid_1d(E, 256 * 16) : 8192, 256,
                                   0, stream>>>(E, d_ed, d_bad);
        GR_HIP(hipGetLastError());
        GR_HIP(hipMemcpyAsync(&h_bad, d_bad, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
        GR_HIP(hipStreamSynchronize(stream));
        (void)hipFree(d_bad);
        inert_ok = h_bad == 0;
      }
    }
    if constexpr (kDelta) {
      // whole 32-entry groups for every bitmap word (k_cand_to_bits)
      GR_TRY(dalloc(&d_cand, (int64_t)n_words * 32));
      k_fill<G><<<1024, 256, 0, stream>>>((int64_t)n_words * 32, P::gatherZero, d_cand);
      GR_HIP(hipGetLastError());
    }
    GR_HIP(hipStreamSynchronize(stream));
    n_active = 0;
    w_edges = 0;
    n_applies = 0;
    iteration = 0;
    delta_chain = delta_ready = false;
    fuse_launched = fused_pending = collect_pending = dense_pending = false;
    fab_launched = fab_pending = last_gather_blocked = false;
    it_gather = it_apply = it_scatter = false;
    bound = true;
    return GR_OK;
  }

  // Stable order of the edges by (primary, secondary) key: a counting sort by
  // the secondary key, then a stable one by the primary key.
  static void edge_order(int32_t nV, int64_t nE, const int32_t* primary,
                         const int32_t* secondary, std::vector<int32_t>& order,
                         std::vector<int32_t>& offsets) {
    std::vector<int32_t> tmp((size_t)nE);
    std::vector<int64_t> next((size_t)nV + 1);
    auto pass = [&](const int32_t* key, const int32_t* in, int32_t* out) {
      std::fill(next.begin(), next.end(), 0);
      for (int64_t i = 0; i < nE; ++i) ++next[(size_t)key[in ? in[i] : i] + 1];
      for (int32_t v = 0; v < nV; ++v) next[(size_t)v + 1] += next[v];
      offsets.assign(next.begin(), next.end());
      for (int64_t i = 0; i < nE; ++i) {
        const int32_t e = in ? in[i] : (int32_t)i;
        out[next[key[e]]++] = e;
      }
    };
    order.resize((size_t)nE);
    pass(secondary, nullptr, tmp.data());
    pass(primary, tmp.data(), order.data());
  }

  // Vertex ranges of at most max_edges edges (a vertex with more is a shard
  // of its own); bounds[s] .. bounds[s + 1].
  static std::vector<int32_t> shard_bounds(const std::vector<int32_t>& off, int64_t max_edges) {
    std::vector<int32_t> b{0};
    const int32_t nV = (int32_t)off.size() - 1;
    int64_t start = 0;
    for (int32_t v = 0; v < nV; ++v) {
      if ((int64_t)off[v + 1] - start > max_edges && v > b.back()) {
        b.push_back(v);
        start = off[v];
      }
    }
    if (b.back() != nV || nV == 0) b.push_back(nV);
    return b;
  }

  int bind_streaming(int32_t nV, void* vdata, size_t velem, int32_t nE, const void* edata,
                     size_t eelem, const int32_t* srcs, const int32_t* dsts) override {
    if constexpr (kScatter) {
      return set_error(GR_ERR_UNSUPPORTED,
                       "streaming: programs whose scatter changes edge data are not supported");
    } else {
      if (nV < 0 || nE < 0 || (nE > 0 && (!srcs || !dsts)))
        return set_error(GR_ERR_INVALID, "setGraph: bad sizes or edge arrays");
      for (int64_t i = 0; i < nE; ++i)
        if (srcs[i] < 0 || srcs[i] >= nV || dsts[i] < 0 || dsts[i] >= nV)
          return set_errorf(GR_ERR_INVALID, "setGraph: edge %lld has a vertex outside [0, %d)",
                            (long long)i, nV);
      // CSC by (dst, src) and CSR by (src, dst), as the device builder orders them
      std::vector<int32_t> csc_idx, csr_idx, off_in, off_out;
      edge_order(nV, nE, dsts, srcs, csc_idx, off_in);
      edge_order(nV, nE, srcs, dsts, csr_idx, off_out);
      gr_graph* gg = new (std::nothrow) gr_graph;
      if (!gg) return set_error(GR_ERR_NOMEM, "streaming: out of host memory");
      gg->nV = nV;
      gg->nE = nE;
      gg->h_off_in = off_in;
      gg->h_off_out = off_out;
      if (hipMalloc((void**)&gg->off_in, sizeof(int32_t) * ((size_t)nV + 1)) != hipSuccess ||
          hipMalloc((void**)&gg->off_out, sizeof(int32_t) * ((size_t)nV + 1)) != hipSuccess) {
        graph_free(gg);
        return set_error(GR_ERR_NOMEM, "streaming: offsets");
      }
      if (hipMemcpy(gg->off_in, off_in.data(), sizeof(int32_t) * ((size_t)nV + 1),
                    hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(gg->off_out, off_out.data(), sizeof(int32_t) * ((size_t)nV + 1),
                    hipMemcpyHostToDevice) != hipSuccess) {
        graph_free(gg);
        return set_error(GR_ERR_HIP, "streaming: offsets copy");
      }
      GR_TRY(bind(gg, true, vdata, velem, edata, eelem, false, true));
      const int rc = stream_buffers(nE, edata, srcs, dsts, csc_idx, csr_idx, off_in, off_out);
      if (rc != GR_OK) {  // leave no half-built engine behind
        release();
        bound = false;
      }
      return rc;
    }
  }

  // Pinned host edge arrays, shard tables and device staging of the streaming
  // mode (after bind() has set up the vertex-sized state).
  int stream_buffers(int32_t nE, const void* edata, const int32_t* srcs, const int32_t* dsts,
                     std::vector<int32_t>& csc_idx, const std::vector<int32_t>& csr_idx,
                     const std::vector<int32_t>& off_in, const std::vector<int32_t>& off_out) {
    {
      const int32_t nV = g->nV;
      // edge arrays into pinned host memory, in CSC (sources, edge data) and
      // CSR (targets) order
      const size_t E = (size_t)nE;
      GR_HIP(hipHostMalloc((void**)&h_srcs, sizeof(int32_t) * (E + 1), hipHostMallocDefault));
      GR_HIP(hipHostMalloc((void**)&h_dsts, sizeof(int32_t) * (E + 1), hipHostMallocDefault));
      for (size_t k = 0; k < E; ++k) {
        h_srcs[k] = srcs[csc_idx[k]];
        h_dsts[k] = dsts[csr_idx[k]];
      }
      if constexpr (!kEdgeEmpty) {
        GR_HIP(hipHostMalloc((void**)&h_ed, sizeof(ED) * (E + 1), hipHostMallocDefault));
        const ED* ed = static_cast<const ED*>(edata);
        for (size_t k = 0; k < E; ++k) h_ed[k] = ed[csc_idx[k]];
      }
      h_csc_eid.swap(csc_idx);
      csc_vb = shard_bounds(off_in, stream_max_edges);
      csr_vb = shard_bounds(off_out, stream_max_edges);
      int64_t max_in = 0, max_out = 0;
      for (size_t s = 0; s + 1 < csc_vb.size(); ++s)
        max_in = std::max<int64_t>(max_in, off_in[csc_vb[s + 1]] - off_in[csc_vb[s]]);
      for (size_t s = 0; s + 1 < csr_vb.size(); ++s)
        max_out = std::max<int64_t>(max_out, off_out[csr_vb[s + 1]] - off_out[csr_vb[s]]);
      const int n_in = (int)csc_vb.size(), n_out = (int)csr_vb.size();
      GR_TRY(dalloc(&d_csc_vb, n_in));
      GR_TRY(dalloc(&d_csr_vb, n_out));
      GR_TRY(dalloc(&d_spos, n_in));
      GR_TRY(dalloc(&d_swsum, n_in));
      GR_TRY(dalloc(&d_scnt, n_out));
      GR_HIP(hipHostMalloc((void**)&h_sbuf, sizeof(int32_t) * (2 * (size_t)n_in + n_out),
                           hipHostMallocDefault));
      GR_HIP(hipMemcpy(d_csc_vb, csc_vb.data(), sizeof(int32_t) * n_in, hipMemcpyHostToDevice));
      GR_HIP(hipMemcpy(d_csr_vb, csr_vb.data(), sizeof(int32_t) * n_out, hipMemcpyHostToDevice));
      GR_TRY(dalloc(&d_escan_sub, (int64_t)nV + 1));
      GR_TRY(dalloc(&d_cbase_sub, (int64_t)nV + 1));
      GR_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
      for (int b = 0; b < 2; ++b) {
        GR_TRY(dalloc(&d_stage_src[b], max_in));
        if constexpr (!kEdgeEmpty) GR_TRY(dalloc(&d_stage_ed[b], max_in));
        GR_TRY(dalloc(&d_stage_dst[b], max_out));
        GR_HIP(hipEventCreateWithFlags(&ev_copied[b], hipEventDisableTiming));
        GR_HIP(hipEventCreateWithFlags(&ev_used[b], hipEventDisableTiming));
        GR_HIP(hipEventRecord(ev_used[b], stream));
      }
      GR_HIP(hipStreamSynchronize(stream));
      return GR_OK;
    }
  }

  // Stages host edge range [e0, e1) of `src` into `dst` on the copy stream
  // once buffer b is free, and makes the compute stream wait for it.
  template <class T>
  int stage(int b, T* dst, const T* src, int64_t e0, int64_t e1) {
    GR_HIP(hipStreamWaitEvent(copy_stream, ev_used[b], 0));
    if (e1 > e0)
      GR_HIP(hipMemcpyAsync(dst, src + e0, sizeof(T) * (size_t)(e1 - e0), hipMemcpyHostToDevice,
                            copy_stream));
    GR_HIP(hipEventRecord(ev_copied[b], copy_stream));
    GR_HIP(hipStreamWaitEvent(stream, ev_copied[b], 0));
    return GR_OK;
  }

  // Gather over the CSC shards that hold active rows; shard k + 1 is copied
  // while shard k is gathered (two staging buffers).
  int gather_streamed() {
    GR_TRY(ensure_list());
    const int n = n_active;
    const int S = (int)csc_vb.size() - 1;
    k_shard_positions<<<grid_1d((int64_t)S + 1, 256), 256, 0, stream>>>(
        S + 1, d_csc_vb, d_active, n, d_escan, d_spos, d_swsum);
    GR_HIP(hipGetLastError());
    GR_HIP(hipMemcpyAsync(h_sbuf, d_spos, sizeof(int32_t) * (S + 1), hipMemcpyDeviceToHost,
                          stream));
    GR_HIP(hipMemcpyAsync(h_sbuf + S + 1, d_swsum, sizeof(int32_t) * (S + 1),
                          hipMemcpyDeviceToHost, stream));
    GR_HIP(hipStreamSynchronize(stream));
    std::vector<int32_t> pos(h_sbuf, h_sbuf + S + 1), wsum(h_sbuf + S + 1, h_sbuf + 2 * S + 2);
    tick();
    int used = 0;
    const int tile = kNT * gather_vt;
    for (int s = 0; s < S; ++s) {
      const int a0 = pos[s], a1 = pos[s + 1];
      if (a1 == a0) continue;
      const int ns = a1 - a0, ws = wsum[s + 1] - wsum[s];
      const int64_t e0 = g->h_off_in[csc_vb[s]], e1 = g->h_off_in[csc_vb[s + 1]];
      const int b = used++ & 1;
      GR_TRY(stage(b, d_stage_src[b], h_srcs, e0, ws ? e1 : e0));
      if constexpr (!kEdgeEmpty) GR_TRY(stage(b, d_stage_ed[b], h_ed, e0, ws ? e1 : e0));
      k_shard_rebase<<<grid_1d((int64_t)ns + 1, 256), 256, 0, stream>>>(
          ns, d_escan, d_cbase, a0, (int)e0, d_escan_sub, d_cbase_sub);
      const int n_tiles = (int)(((int64_t)ns + ws + tile - 1) / tile);
      k_merge_partition<<<grid_1d((int64_t)n_tiles + 1, 256), 256, 0, stream>>>(
          d_escan_sub, ns, ws, n_tiles, tile, d_part);
      const ED* sed = kEdgeEmpty ? reinterpret_cast<const ED*>(d_gtmp) : d_stage_ed[b];
#define GR_STREAM_GATHER(VT)                                                                   \
  k_gather_tiles<P, VT><<<n_tiles, kNT, 0, stream>>>(ns, ws, d_active + a0, d_escan_sub, d_part, \
                                                     d_cbase_sub, d_stage_src[b], d_vd, sed,    \
                                                     d_gtmp + a0, d_carry, d_head, nullptr,     \
                                                     d_srcval, nullptr)
      switch (gather_vt) {
        case 4: GR_STREAM_GATHER(4); break;
        case 16: GR_STREAM_GATHER(16); break;
        default: GR_STREAM_GATHER(8); break;
      }
#undef GR_STREAM_GATHER
      k_gather_fixup<P><<<grid_1d(n_tiles, 256), 256, 0, stream>>>(n_tiles, d_carry, d_head,
                                                                   d_gtmp + a0);
      GR_HIP(hipGetLastError());
      GR_HIP(hipEventRecord(ev_used[b], stream));
      if (ws) {
        stats.streamed_bytes += (e1 - e0) * (int64_t)(4 + (kEdgeEmpty ? 0 : sizeof(ED)));
        stats.streamed_shards += 1;
      }
    }
    tock(kPhGather);
    stats.active_vertices += n;
    stats.gathered_edges += w_edges;
    return GR_OK;
  }

  // Activation over the CSR shards that hold changed vertices.
  int scatter_streamed() {
    const int S = (int)csr_vb.size() - 1;
    if (S > 0) {
      k_shard_changed<<<S, 256, 0, stream>>>(S, d_csr_vb, d_chg_bits, d_scnt);
      GR_HIP(hipGetLastError());
      GR_HIP(hipMemcpyAsync(h_sbuf, d_scnt, sizeof(int32_t) * S, hipMemcpyDeviceToHost, stream));
      GR_HIP(hipStreamSynchronize(stream));
    }
    std::vector<int32_t> cnt(h_sbuf, h_sbuf + S);
    tick();
    int used = 0;
    for (int s = 0; s < S; ++s) {
      if (!cnt[s]) continue;
      const int u0 = csr_vb[s], u1 = csr_vb[s + 1];
      const int64_t e0 = g->h_off_out[u0], e1 = g->h_off_out[u1];
      if (e1 == e0) continue;
      const int b = used++ & 1;
      GR_TRY(stage(b, d_stage_dst[b], h_dsts, e0, e1));
      const unsigned grid = std::min<unsigned>(grid_1d((int64_t)u1 - u0, 256), 8192);
      k_shard_push<<<grid, 256, 0, stream>>>(u0, u1, (int)e0, g->off_out, d_stage_dst[b],
                                             d_chg_bits, d_bitmap);
      GR_HIP(hipGetLastError());
      GR_HIP(hipEventRecord(ev_used[b], stream));
      stats.streamed_bytes += (e1 - e0) * 4;
      stats.streamed_shards += 1;
    }
    tock(kPhScatter);
    return GR_OK;
  }

  static_assert(sizeof(VD) % 4 == 0, "vertex data is moved in 4-byte words");
  // user order -> layout order: out[i] = in[perm[i]]
  void to_layout(const VD* in, VD* out) {
    k_permute_words<<<grid_1d(g->nV, 256), 256, 0, stream>>>(
        reinterpret_cast<const uint32_t*>(in), d_perm, g->nV, (int)(sizeof(VD) / 4),
        reinterpret_cast<uint32_t*>(out));
  }
  // Device pointer to the vertex data in user order (d_vd itself without a
  // reordered layout).
  const VD* user_order_vd() {
    if (!d_inv) return d_vd;
    k_permute_words<<<grid_1d(g->nV, 256), 256, 0, stream>>>(
        reinterpret_cast<const uint32_t*>(d_vd), d_inv, g->nV, (int)(sizeof(VD) / 4),
        reinterpret_cast<uint32_t*>(d_vd_user));
    return d_vd_user;
  }

  const ED* edge_csr_ptr() const {
    if constexpr (kEdgeEmpty)
      return reinterpret_cast<const ED*>(d_gtmp);  // never dereferenced
    else
      return d_ed_csr;
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
    GR_TRY(drop_candidates());
    const int n = end - start;
    if (d_inv && n < g->nV) return set_active_list(start, end);
    GR_TRY(drop_hub_marks());
    src_flags_ok = gather_flags = false;  // the changed bitmaps are cleared below
    GR_HIP(hipMemsetAsync(d_bitmap, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_bits, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_prev, 0, sizeof(uint32_t) * (n_words + 1), stream));
    // applies since the last nextIter belong to the replaced state
    GR_HIP(hipMemsetAsync(d_applied, 0,
                          sizeof(unsigned long long) * kAppliedSlots * kAppliedStride, stream));
    // An all-active iteration gathers every in-edge of every vertex, which
    // establishes the delta-gather invariant; a partial range does not.
    delta_chain = (start == 0 && end == g->nV);
    delta_ready = false;
    it_gather = it_apply = it_scatter = false;
    k_set_active_range<<<grid_1d((int64_t)n + 1, 256), 256, 0, stream>>>(start, n, g->off_in,
                                                                         d_active, d_escan, d_cbase,
                                                                         d_ctr);
    GR_HIP(hipGetLastError());
    list_valid = true;
    n_active = n;
    w_edges = g->h_off_in[end] - g->h_off_in[start];
    n_applies = 0;
    iteration = 0;
    return GR_OK;
  }

  // setActive on a sub-range of user ids with the ordered layout: the layout
  // ids of the range, ascending, with their in-edge scan built on the host.
  int set_active_list(int32_t start, int32_t end) {
    GR_TRY(drop_candidates());
    const int n = end - start;
    std::vector<int32_t> ids((size_t)n), escan((size_t)n + 1), cbase((size_t)n + 1);
    if (n > 0)
      GR_HIP(hipMemcpy(ids.data(), d_inv + start, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    std::sort(ids.begin(), ids.end());
    int64_t acc = 0;
    for (int i = 0; i < n; ++i) {
      escan[i] = (int32_t)acc;
      cbase[i] = g->h_off_in[ids[i]] - (int32_t)acc;
      acc += g->h_off_in[ids[i] + 1] - g->h_off_in[ids[i]];
    }
    escan[n] = (int32_t)acc;
    cbase[n] = 0;
    GR_TRY(drop_hub_marks());
    src_flags_ok = gather_flags = false;  // the changed bitmaps are cleared below
    GR_HIP(hipMemsetAsync(d_bitmap, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_bits, 0, sizeof(uint32_t) * (n_words + 1), stream));
    GR_HIP(hipMemsetAsync(d_chg_prev, 0, sizeof(uint32_t) * (n_words + 1), stream));
    if (n > 0) {
      GR_HIP(hipMemcpyAsync(d_active, ids.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                            stream));
      GR_HIP(hipMemcpyAsync(d_cbase, cbase.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                            stream));
    }
    GR_HIP(hipMemcpyAsync(d_escan, escan.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice,
                          stream));
    GR_HIP(hipMemsetAsync(d_applied, 0,
                          sizeof(unsigned long long) * kAppliedSlots * kAppliedStride, stream));
    k_set_counters<<<1, 1, 0, stream>>>(d_ctr, n, (int32_t)acc);
    GR_HIP(hipGetLastError());
    list_valid = true;
    GR_HIP(hipStreamSynchronize(stream));  // the host vectors go out of scope
    n_active = n;
    w_edges = (int32_t)acc;
    n_applies = 0;
    iteration = 0;
    delta_chain = false;
    delta_ready = false;
    it_gather = it_apply = it_scatter = false;
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
    it_gather = true;
    GR_TRY(collect_flush());  // a previous gather whose apply did not run
    dense_pending = false;
    last_gather_blocked = false;
    if (fab_pending) {  // computed by the previous scatterActivate
      fab_pending = false;
      if (P::kHasGather && n_active > 0) {
        dense_pending = true;
        last_gather_blocked = true;
        stats.active_vertices += n_active;
        stats.gathered_edges += w_edges;
        stats.tile_gather_rows += n_active;
        stats.tile_gather_edges += w_edges;
        stats.blocked_gathers += 1;
        stats.fused_activations += 1;
        if (kDelta) stats.delta_gathers += 1;
      }
      return GR_OK;
    }
    if (!P::kHasGather || n_active == 0) return GR_OK;
    if (streaming) return gather_streamed();
    const int n = n_active, w = w_edges;
    const uint32_t* dbits = nullptr;
    if constexpr (kDelta) {
      if (fused_pending && delta_ready) {
        fused_pending = false;
        return gather_collect();
      }
      if (delta_ready) {
        if (delta_mode == 3 || (delta_mode == 1 && changed_share < push_share))
          return gather_push();
        if (delta_mode == 2) dbits = d_chg_prev;
        // the blocked gather reads the changed bitmap in source order, at
        // little cost, and skips the unchanged sources' values and edge data
        if (delta_mode == 1 && blocked_gather_applies(w)) dbits = d_chg_prev;
      } else if (kInert && inert_ok && delta_mode != 0) {
        // The non-inert vertices stand in for the changed set: the gather's
        // values are those of the full gather, so an all-active start still
        // establishes the delta invariant.  Pushing needs every vertex active
        // (a pushed candidate is collected and reset by its active row only).
        GR_TRY(inert_scan());
        const double share = w > 0 ? (double)chg_edges / (double)w : 1.0;
        if (n == g->nV && (delta_mode == 3 || (delta_mode == 1 && share < push_share)))
          return gather_push();
        if (delta_mode != 1 || share < push_share || blocked_gather_applies(w))
          dbits = d_chg_prev;
      }
    }
    if (blocked_gather_applies(w)) return gather_blocked(n, w, dbits);
    GR_TRY(ensure_list());
    const int tile = kNT * gather_vt;
    const int n_tiles = (int)(((int64_t)n + w + tile - 1) / tile);
    k_merge_partition<<<grid_1d((int64_t)n_tiles + 1, 256), 256, 0, stream>>>(d_escan, n, w,
                                                                              n_tiles, tile, d_part);
    tick();
#define GR_GATHER_LAUNCH(VT)                                                                \
  k_gather_tiles<P, VT><<<n_tiles, kNT, 0, stream>>>(n, w, d_active, d_escan, d_part, d_cbase, \
                                                     g->srcs, d_vd, edge_ptr(), d_gtmp,       \
                                                     d_carry, d_head, dbits, d_srcval, g->csc_eid)
    switch (gather_vt) {
      case 4: GR_GATHER_LAUNCH(4); break;
      case 16: GR_GATHER_LAUNCH(16); break;
      default: GR_GATHER_LAUNCH(8); break;
    }
#undef GR_GATHER_LAUNCH
    tock(kPhGather);
    k_gather_fixup<P><<<grid_1d(n_tiles, 256), 256, 0, stream>>>(n_tiles, d_carry, d_head, d_gtmp);
    GR_HIP(hipGetLastError());
    stats.active_vertices += n;
    stats.gathered_edges += w;
    stats.tile_gather_rows += n;
    stats.tile_gather_edges += w;
    if (dbits && delta_ready) stats.delta_gathers += 1;
    return GR_OK;
  }

  bool blocked_gather_applies(int w) const {
    return kBlocked && blocked_share > 0 && g->nE > 0 &&
           (double)w >= blocked_share * (double)g->nE;
  }

  // The graph's source-listed row blocks, the edge data in their order and
  // d_gdense, set up on the first blocked gather of a binding.
  int blocked_setup() {
    if constexpr (kBlocked) {
      if (!blk) {
        blk_rows = kBlockedSum ? kBlockRowsWide : graph_skewed(g) ? kMinBlockRows : kMinBlockRowsFlat;
        GR_TRY(graph_blocked(g, blk_rows, &blk));
        GR_TRY(dalloc(&d_gdense, g->nV));
        k_fill<G><<<1024, 256, 0, stream>>>(g->nV, P::gatherZero, d_gdense);
        GR_HIP(hipGetLastError());
        if constexpr (!kEdgeEmpty) {
          static_assert(sizeof(ED) % 4 == 0, "edge data is moved as 4-byte words");
          GR_TRY(dalloc(&d_ed_blk, blk->n_slots));
          k_permute_words<<<grid_1d(blk->n_slots, 256), 256, 0, stream>>>(
              reinterpret_cast<const uint32_t*>(d_ed), blk->bpos, blk->n_slots,
              (int)(sizeof(ED) / 4), reinterpret_cast<uint32_t*>(d_ed_blk));
          GR_HIP(hipGetLastError());
        }
      }
    }
    return GR_OK;
  }

  // k_gather_blocked over the row blocks, for the blocks' row limit.
  void launch_blocked_min(const uint32_t* dbits, uint32_t* frontier, DevCounters* ctr,
                          int mode = 0, long long thr = 0, long long fuse_max = -1) {
    if constexpr (kBlockedMin) {
      if (blk_rows == kBlockRows)
        k_gather_blocked<P, kBlockRows><<<blk->n_blocks, kBlkNT, 0, stream>>>(
            blk->brow, blk->bnr, blk->boff, blk->bpk, blk->gbase, d_ed_blk, d_vd, d_srcval,
            dbits, d_gdense, frontier, ctr, mode, thr, fuse_max);
      else
        k_gather_blocked<P, kBlockRowsWide><<<blk->n_blocks, kBlkNT, 0, stream>>>(
            blk->brow, blk->bnr, blk->boff, blk->bpk, blk->gbase, d_ed_blk, d_vd, d_srcval,
            dbits, d_gdense, frontier, ctr, mode, thr, fuse_max);
    } else {
      (void)dbits;
      (void)frontier;
      (void)ctr;
      (void)mode;
      (void)thr;
      (void)fuse_max;
    }
  }

  // Gather of every row through the graph's source-listed row blocks
  // (k_gather_blocked); apply reads the active rows' values from d_gdense.
  int gather_blocked(int n, int w, const uint32_t* dbits) {
    if constexpr (kBlocked) {
      GR_TRY(blocked_setup());
      tick();
      if (blk->n_blocks > 0) {
        if constexpr (kBlockedSum)
          k_gather_blocked_sum<P><<<blk->n_blocks, kBlkNT, 0, stream>>>(
              blk->brow, blk->bnr, blk->boff, blk->bpk, blk->gbase, d_ed_blk, d_vd, d_srcval,
              d_gdense, nullptr, nullptr, nullptr);
        else
          launch_blocked_min(dbits, nullptr, nullptr);
      }
      tock(kPhGather);
      GR_HIP(hipGetLastError());
      dense_pending = true;
      last_gather_blocked = true;
      stats.active_vertices += n;
      stats.gathered_edges += w;
      stats.tile_gather_rows += n;
      stats.tile_gather_edges += w;
      stats.blocked_gathers += 1;
      if (dbits && delta_ready) stats.delta_gathers += 1;
    } else {
      (void)n;
      (void)w;
      (void)dbits;
    }
    return GR_OK;
  }

  // The next gather after a fused scatterActivate: the candidates are already
  // relaxed; read and reset them for the active rows.
  int gather_collect() {
    const int n = n_active;
    collect_pending = true;  // apply reads the candidates
    stats.active_vertices += n;
    stats.gathered_edges += w_edges;
    stats.delta_gathers += 1;
    stats.push_gathers += 1;
    stats.push_edges += chg_edges;
    stats.fused_gathers += 1;
    stats.fused_edges += chg_edges;
    return GR_OK;
  }

  // Candidates of a push gather whose apply did not run: reset them (the
  // values go to d_gtmp, unused).
  int collect_flush() {
    if constexpr (kDelta) {
      if (collect_pending) {
        GR_TRY(ensure_list());
        k_relax_collect<P><<<grid_1d(n_active, 256), 256, 0, stream>>>(n_active, d_active, d_cand,
                                                                       d_gtmp);
        GR_HIP(hipGetLastError());
      }
    }
    collect_pending = false;
    return GR_OK;
  }

  // Candidates relaxed by a fused scatterActivate whose gather will not run
  // (the active set was replaced or the vertex data reset): back to gatherZero.
  // Hub marks of a push whose frontier is being cleared (setActive): the
  // next compaction must not fold them in.
  int drop_hub_marks() {
    if (hub.flags) {
      GR_HIP(hipMemsetAsync(d_hub_flags, 0, (size_t)n_hub, stream));
      hub = HubFold();
    }
    return GR_OK;
  }

  int drop_candidates() {
    GR_TRY(ensure_chg_list());  // before the changed bitmaps are cleared
    GR_TRY(collect_flush());
    fab_pending = fab_launched = false;
    last_gather_blocked = false;
    if constexpr (kDelta) {
      if (fused_pending) {
        k_fill<G><<<1024, 256, 0, stream>>>(g->nV, P::gatherZero, d_cand);
        GR_HIP(hipGetLastError());
      }
    }
    fused_pending = false;
    return GR_OK;
  }

  // Bitmap (in d_chg_prev) and list (d_chg_v / d_chg_e, chg_count / chg_edges)
  // of the non-inert vertices.
  int inert_scan() {
    if constexpr (kInert) {
      chg_pending = chg_cond = false;  // the list is replaced by the non-inert vertices
      tick();
      GR_HIP(hipMemsetAsync(&d_ctr->inert, 0, sizeof(unsigned long long), stream));
      k_inert_scan<P><<<grid_1d(g->nV, kApplyChunk), kApplyNT, 0, stream>>>(
          g->nV, n_words, d_vd, g->off_out, d_chg_v, d_chg_e, d_chg_prev, d_ctr);
      GR_HIP(hipGetLastError());
      GR_HIP(hipMemcpyAsync(&h_ctr->inert, &d_ctr->inert, sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stream));
      tock(kPhInert);
      GR_HIP(hipStreamSynchronize(stream));
      chg_count = (int64_t)(h_ctr->inert >> 32);
      chg_edges = (int64_t)(uint32_t)h_ctr->inert;
      stats.inert_gathers += 1;
    }
    return GR_OK;
  }

  // Delta gather by push: relax the out-edges of the previous iteration's
  // changed list into d_cand, then read the candidates of the active rows.
  int gather_push() {
    GR_TRY(ensure_chg_list());
    const int n = n_active;
    const int n_c = (int)chg_count, w_c = (int)chg_edges;
    tick();
    if (w_c > 0) {
      const int n_tiles = (w_c + kTile - 1) / kTile;
      k_relax_partition<<<grid_1d((int64_t)n_tiles + 1, 256), 256, 0, stream>>>(n_c, w_c, d_chg_e,
                                                                                d_push_part);
      const int grid = n_tiles < 8192 ? n_tiles : 8192;
      k_relax_tiles<P><<<grid, kNT, 0, stream>>>(w_c, nullptr, -1, d_push_part, d_chg_v, d_chg_e,
                                                 g->off_out, g->dsts, d_vd, edge_csr_ptr(),
                                                 d_cand, nullptr, 0);
    }
    tock(kPhPush);
    GR_HIP(hipGetLastError());
    collect_pending = true;  // apply reads the candidates
    stats.active_vertices += n;
    stats.gathered_edges += w_edges;
    if (delta_ready) stats.delta_gathers += 1;  // else the list holds the non-inert vertices
    stats.push_gathers += 1;
    stats.push_edges += w_c;
    return GR_OK;
  }

  // The changed bitmap feeds pull activation and the next delta gather.
  bool changed_bits_on() const { return dir_mode != 1 || kDelta || streaming; }

  int apply() override {
    it_apply = true;
    if (pre_applied) {  // queued by nextIter (run loop): the same launch
      pre_applied = false;
      if (n_active == 0) return GR_OK;
      ++n_applies;
      collect_pending = false;
      dense_pending = false;
      chg_pending = true;
      return GR_OK;
    }
    if (n_active == 0) return GR_OK;
    ++n_applies;
    const int use_gather = !P::kHasGather ? 0 : collect_pending ? 2 : dense_pending ? 3 : 1;
    collect_pending = false;
    dense_pending = false;
    // from the active bitmap unless the list is there anyway or the gather
    // values are in list order (use_gather 1)
    const bool bits = !list_valid && use_gather != 1 &&
                      (int64_t)n_active * kBitsApplyDiv >= (int64_t)g->nV;
    if (!bits) GR_TRY(ensure_list());
    // a second apply of the same iteration appends to the first one's list
    GR_TRY(ensure_chg_list());
    tick();
    uint32_t* cbits = changed_bits_on() ? d_chg_bits : nullptr;
    G* cv = use_gather == 3 ? d_gdense : d_cand;
    // The first apply of an iteration over the bitmap defers its changed
    // list (no per-workgroup atomic on the list counter); the changed bitmap
    // holds exactly its vertices then, from which k_chg_write makes the list.
    const bool defer = bits && cbits && n_applies == 1;
    ApplyDefer ad;
    if (defer) {
      ad.cc = d_cc;
      ad.sc = d_sc;
      ad.sc_final = d_sc_final;
    }
    const uint32_t* fprev = nullptr;
    GR_TRY(source_flags_for_apply(defer, &fprev));
    if (bits)
      k_apply<P, true><<<grid_1d(g->nV, kApplyChunk), kApplyNT, 0, stream>>>(
          g->nV, nullptr, d_act_bits, d_gtmp, use_gather, d_vd, g->off_out, d_deg_out8, iteration, d_chg_v,
          d_chg_e, cbits, d_ctr, d_srcval, cv, d_applied, ad, fprev);
    else
      k_apply<P, false><<<grid_1d(n_active, kApplyChunk), kApplyNT, 0, stream>>>(
          n_active, d_active, nullptr, d_gtmp, use_gather, d_vd, g->off_out, d_deg_out8, iteration, d_chg_v,
          d_chg_e, cbits, d_ctr, d_srcval, cv, d_applied, ApplyDefer());
    if (defer) chg_pending = true;
    tock(kPhApply);
    GR_HIP(hipGetLastError());
    return GR_OK;
  }

  int end_run() override {
    pre_applied = false;
    return GR_OK;
  }

  // Source flags for an apply: a flag-keeping apply (the first of an
  // iteration, bitmap form, with changed bitmaps) clears the previous
  // apply's flags that it does not set again, which needs the flags to be
  // at most that apply's changed set (d_chg_prev): otherwise the kept values
  // are rewritten clean first.  Any other apply leaves flags behind that the
  // fused blocked sum must not read.
  int source_flags_for_apply(bool keeps, const uint32_t** fprev) {
    *fprev = nullptr;
    if constexpr (source_flag_v<P>) {
      if (keeps && changed_bits_on() && !streaming) {
        if (!src_flags_ok) GR_TRY(refresh_source_values());
        *fprev = d_chg_prev;
        src_flags_ok = gather_flags = true;
      } else {
        src_flags_ok = gather_flags = false;
      }
    } else {
      (void)keeps;
    }
    return GR_OK;
  }

  // The changed list of a deferred apply, from the changed bitmap it left.
  // With `standdown` (scatterActivate launching the blocked gather in place
  // of the pull, mode / thr / fuse_max its rule) the kernel writes nothing
  // when that gather runs; a later phase of the same iteration that needs the
  // list then writes it unconditionally (chg_cond).
  int ensure_chg_list(bool standdown = false, int mode = 0, long long thr = 0,
                      long long fuse_max = -1) {
    if (!chg_pending && !chg_cond) return GR_OK;
    standdown = standdown && chg_pending;
    chg_pending = false;
    chg_cond = standdown;
    const int grid = n_apply_chunks < 8192 ? n_apply_chunks : 8192;
    k_chg_write<<<grid < 1 ? 1 : grid, kChgNT, 0, stream>>>(
        d_chg_bits, g->nV, g->off_out, d_deg_out8, d_cc, d_sc_final, n_apply_chunks, d_chg_v,
        d_chg_e, d_ctr,
        standdown ? 1 : 0, mode, thr, fuse_max);
    GR_HIP(hipGetLastError());
    return GR_OK;
  }

  // Active sets of at least nV / kBitsApplyDiv vertices are applied from the
  // bitmap: a pass over all chunks costs about what writing and reading the
  // list saves (BFS at RMAT-24: 1.4 % active levels ran faster from the list).
  static constexpr int kBitsApplyDiv = 32;

  // At least the out-edges of any k vertices: the largest buckets first, the
  // last one taken at its bucket's largest degree.
  int64_t top_out_edges(int64_t k) const {
    int64_t sum = 0;
    for (int b = 31; b >= 0 && k > 0; --b) {
      if (k >= deg_cnt[b]) {
        sum += deg_sum[b];
        k -= deg_cnt[b];
      } else {
        sum += k * (((int64_t)2 << b) - 1);
        k = 0;
      }
    }
    return sum;
  }

  long long pull_threshold() const {
    return (long long)((double)g->nE / (dir_alpha > 0 ? dir_alpha : 1.0));
  }

  int scatter_activate(bool have_scatter) override {
    it_scatter = true;
    GR_TRY(collect_flush());  // a push gather whose apply did not run
    if (streaming) return scatter_streamed();
    if constexpr (kBlockedSum) {
      // after a dense (blocked) gather the next one is most likely dense too:
      // gather every row now and mark the pull's activations in the same pass
      if (last_gather_blocked && blk && it_gather && it_apply && changed_bits_on() &&
          blocked_share > 0) {
        tick();
        if (blk->n_blocks > 0)
          k_gather_blocked_sum<P><<<blk->n_blocks, kBlkNT, 0, stream>>>(
              blk->brow, blk->bnr, blk->boff, blk->bpk, blk->gbase, d_ed_blk, d_vd, d_srcval,
              d_gdense, d_chg_bits, d_bitmap, d_ctr, gather_flags ? 1 : 0);
        tock(kPhGather);
        GR_HIP(hipGetLastError());
        fab_pending = true;
        last_gather_blocked = false;
        return GR_OK;
      }
    }
    const long long thr = pull_threshold();
    // Fused activation + next push gather, decided on the device from the
    // changed out-edges: only while the next gather may use the delta rule
    // (the same conditions next_iter checks) and not in forced-pull mode.
    long long fuse_max = -1;
    if constexpr (kDelta) {
      if (delta_chain && it_gather && it_apply && dir_mode != 2 &&
          (delta_mode == 1 || delta_mode == 3))
        fuse_max = delta_mode == 3 || fuse_share >= 1.0
                       ? (long long)g->nE : (long long)(fuse_share * (double)g->nE);
      if (fuse_share <= 0 && delta_mode != 3) fuse_max = -1;
    }
    fuse_launched = fuse_max >= 0;
    // The changed list holds active vertices only, each apply call since the
    // active set was built appending at most all of them: their out-edges
    // are at most `bound`.  Kernels whose device-side condition that rules
    // out are not launched.
    const long long bound =
        (long long)(top_out_edges(n_active) * (n_applies > 1 ? n_applies : 1));
    const bool may_pull = dir_mode == 2 || (dir_mode == 0 && bound > thr);
    const bool may_push = dir_mode != 2 && bound > 0 && !(fuse_max >= 0 && bound <= fuse_max);
    // tiles the changed out-edges can fill: the push kernels' grids (they
    // stride over the tiles, so any grid of at least one workgroup is whole)
    const int push_tiles = (int)std::min<long long>(max_push_tiles, bound / kTile + 1);
    // (2,048 workgroups fill every CU; a launch the device stands down costs
    // less with fewer of them)
    const int push_grid = push_tiles < 2048 ? (push_tiles < 1 ? 1 : push_tiles) : 2048;
    // Min programs whose next gather may use the delta rule: where the pull
    // would run, the blocked gather of the changed sources' out-edges takes
    // its place and also computes the next gather's values (timed with the
    // gathers).  Its first workgroup records in ctr->fab whether it ran.
    bool fab = false;
    if constexpr (kBlockedMin) {
      fab = may_pull && delta_chain && it_gather && it_apply && delta_mode == 1 &&
            blocked_share > 0 && g->nE > 0;
      if (fab) {
        GR_TRY(blocked_setup());
        fab = blk->n_blocks > 0;
      }
    }
    // The push, relax and scatter kernels below read the changed list; where
    // the pull (or the blocked gather in its place) runs they stand down.
    // After the blocked gather the next gather is that gather's values, and
    // programs without the delta rule never push a gather: the list is then
    // not needed (unless Program::scatter runs).
    {
      bool need_always = false;
      if constexpr (kScatter) need_always = have_scatter;
      const bool may_skip = fab || (!kDelta && may_pull);
      GR_TRY(ensure_chg_list(may_skip && !need_always, dir_mode, thr, fuse_max));
    }
    if constexpr (kBlockedMin) {
      if (fab) {
        tick();
        launch_blocked_min(d_chg_bits, d_bitmap, d_ctr, dir_mode, thr, fuse_max);
        tock(kPhGather);
        GR_HIP(hipGetLastError());
      }
    }
    fab_launched = fab;
    tick();
    if (may_pull && !fab) {
      // (grid-stride over the vertices: a capped grid also stands down fast)
      const unsigned pull_grid = GR_PULL_GRID > 0 && grid_1d(g->nV, 256) > GR_PULL_GRID ? GR_PULL_GRID : grid_1d(g->nV, 256);
      k_pull_stream<<<pull_grid, 256, 0, stream>>>(d_ctr, dir_mode, thr, fuse_max,
                                                              g->nV, g->off_in, g->srcs,
                                                              d_chg_bits, d_bitmap);
      if (n_heavy > 0)
        k_pull_heavy<<<std::min(grid_1d(n_heavy, 4), 4096u), 256, 0, stream>>>(
            d_ctr, dir_mode, thr, fuse_max, d_heavy, n_heavy, g->off_in, g->srcs, d_chg_bits,
            d_bitmap);
    }
    {
      // partition of the changed out-edges (push or fused pass); also
      // records in ctr->fused whether the fused pass runs
      k_push_partition<<<grid_1d((int64_t)push_tiles + 1, 256), 256, 0, stream>>>(
          d_ctr, dir_mode, thr, d_chg_e, d_push_part, max_push_tiles, fuse_max, 1);
      if (may_push) {
        k_push_tiles<<<push_grid, kNT, 0, stream>>>(d_ctr, dir_mode, thr, fuse_max, d_push_part,
                                               d_chg_v, d_chg_e, g->off_out, g->dsts, d_bitmap,
                                               d_hub_flags, n_hub);
        // nextIter's compaction folds the hub marks into the frontier (under
        // the rule the push ran by)
        if (n_hub > 0) {
          hub.flags = d_hub_flags;
          hub.n_hub = n_hub;
          hub.mode = dir_mode;
          hub.threshold = thr;
          hub.fuse_max = fuse_max;
        }
      }
    }
    if constexpr (kScatter) {
      // Program::scatter, as the reference's CPU engine runs it when
      // haveScatter (refgas.h:192-196); its GPU engine ignores the flag
      if (have_scatter) {
        k_push_partition<<<grid_1d((int64_t)max_push_tiles + 1, 256), 256, 0, stream>>>(
            d_ctr, 1, 0, d_chg_e, d_push_part, max_push_tiles, -1, 0);
        const int grid = max_push_tiles < 8192 ? max_push_tiles : 8192;
        k_scatter_tiles<P><<<grid < 1 ? 1 : grid, kNT, 0, stream>>>(
            d_ctr, d_push_part, d_chg_v, d_chg_e, g->off_out, g->dsts, g->csr_eid, d_vd, d_ed);
      }
    }
    tock(kPhScatter);
    if constexpr (kDelta) {
      // the fused pass is the next gather's relax: timed with the push gathers
      if (fuse_launched && bound > 0) {
        tick();
        k_relax_tiles<P><<<push_grid, kNT, 0, stream>>>(0, d_ctr, fuse_max, d_push_part, d_chg_v,
                                                   d_chg_e, g->off_out, g->dsts, d_vd,
                                                   edge_csr_ptr(), d_cand, d_bitmap, 1);
        const int64_t n_scan = (int64_t)n_words * 32;
        if (bound > kScanMinEdges)
          k_cand_to_bits<P><<<grid_1d(n_scan, 256) < 2048 ? grid_1d(n_scan, 256) : 2048, 256, 0,
                              stream>>>(d_ctr, fuse_max, n_words, d_cand, d_bitmap);
        tock(kPhPush);
      }
    }
    GR_HIP(hipGetLastError());
    return GR_OK;
  }

  int next_iter() override {
    GR_TRY(collect_flush());
    const int need_deg = P::kHasGather ? 1 : 0;
    tick();
    ++ctr_seq;
    CompactFinal fin;
    fin.gsum = d_cgsum;
    fin.grp = d_cgrp;
    fin.tot = d_cgtot;
    fin.gdone = d_cgdone;
    fin.ctr = d_ctr;
    fin.escan = need_deg ? d_escan : nullptr;
    fin.mode = dir_mode;
    fin.threshold = pull_threshold();
    fin.host = d_hctr;
    fin.host_seq = d_hseq;
    fin.seq = ctr_seq;
    fin.applied = d_applied;
    k_compact_count<<<n_cblocks, kCompactNT, 0, stream>>>(
        d_bitmap, n_words, g->nV, g->off_in, d_deg_in8, need_deg, d_blk, d_act_bits,
        changed_bits_on() ? d_chg_prev : nullptr, hub, fin);
    hub = HubFold();
    list_valid = false;
    tock(kPhNext);
    GR_HIP(hipGetLastError());
    // a deferred list no phase asked for: the next delta gather may push the
    // out-edges of this iteration's changed vertices, other programs never
    // read it again
    if (kDelta && chg_pending) GR_TRY(ensure_chg_list());
    chg_pending = chg_cond = false;
    // this iteration's changes become the previous ones; the cleared buffer
    // receives the next apply's
    std::swap(d_chg_bits, d_chg_prev);
    delta_chain = delta_chain && it_gather && it_apply && it_scatter;
    delta_ready = kDelta && delta_chain && changed_bits_on();
    it_gather = it_apply = it_scatter = false;
    if constexpr (kBlockedSum) {
      // In the run loop after the fused blocked sum, the next gather only
      // hands its values on and the next apply is the bitmap form over them
      // whatever the frontier's size (an empty one leaves every chunk
      // unread): queue it now, so that the GPU does not wait for the host to
      // learn the size first.
      if (run_mode && fab_pending && P::kHasGather && changed_bits_on() && !streaming) {
        const uint32_t* fprev = nullptr;
        GR_TRY(source_flags_for_apply(true, &fprev));
        tick();
        ApplyDefer ad;
        ad.cc = d_cc;
        ad.sc = d_sc;
        ad.sc_final = d_sc_final;
        k_apply<P, true><<<grid_1d(g->nV, kApplyChunk), kApplyNT, 0, stream>>>(
            g->nV, nullptr, d_act_bits, d_gtmp, 3, d_vd, g->off_out, d_deg_out8, iteration + 1,
            d_chg_v,
            d_chg_e, d_chg_bits, d_ctr, d_srcval, d_gdense, d_applied, ad, fprev);
        tock(kPhApply);
        GR_HIP(hipGetLastError());
        pre_applied = true;
      }
    }
    if constexpr (kBlockedMin) {
      // The same after the blocked gather that scatterActivate launched in
      // place of the pull, when the next gather may use its values: the
      // queued apply runs only if that gather ran on the device (it records
      // that in ctr->fab); the host learns which below.
      if (run_mode && fab_launched && delta_ready && P::kHasGather && !streaming) {
        tick();
        ApplyDefer ad;
        ad.cc = d_cc;
        ad.sc = d_sc;
        ad.sc_final = d_sc_final;
        k_apply<P, true><<<grid_1d(g->nV, kApplyChunk), kApplyNT, 0, stream>>>(
            g->nV, nullptr, d_act_bits, d_gtmp, 3, d_vd, g->off_out, d_deg_out8, iteration + 1,
            d_chg_v, d_chg_e, d_chg_bits, d_ctr, d_srcval, d_gdense, d_applied, ad, nullptr, 1);
        tock(kPhApply);
        GR_HIP(hipGetLastError());
        gated_apply = true;
      }
    }
    GR_TRY(wait_counters(ctr_seq));
    if constexpr (kDelta) {
      const bool fused = fuse_launched && h_ctr->fused;
      fuse_launched = false;
      fused_pending = fused;
      if (fused && !delta_ready) GR_TRY(drop_candidates());  // not reached: same conditions
    }
    if constexpr (kBlockedMin) {
      if (fab_launched && !h_ctr->fab) stats.idle_gather_launches += 1;
      fab_pending = fab_launched && h_ctr->fab && delta_ready;
      fab_launched = false;
      if (gated_apply) {  // it ran exactly where the next apply is its form
        pre_applied = fab_pending;
        gated_apply = false;
      }
    }
    const unsigned long long c = h_ctr->chg_last;
    stats.changed_vertices += (int64_t)h_ctr->applied_last;
    stats.pushed_edges += (int64_t)(uint32_t)c;
    stats.pull_iterations += h_ctr->pull ? 1 : 0;
    if (n_active > 0) stats.iterations += 1;
    n_active = h_ctr->n_active;
    w_edges = h_ctr->w_edges;
    n_applies = 0;
    changed_share = w_edges > 0 ? (double)(uint32_t)c / (double)w_edges : 1.0;
    chg_count = (int64_t)(c >> 32);
    chg_edges = (int64_t)(uint32_t)c;
    ++iteration;
    return GR_OK;
  }

  // Waits until k_compact_count has published counters number `want` to h_ctr.
  // The host polls the sequence number instead of synchronising the stream:
  // it learns the next frontier's size as soon as the scan has published it
  // and queues the next iteration's kernels behind it.  The polling is
  // bounded; after kSpinLimit the stream is synchronised, after which the
  // number must be there.
  int wait_counters(uint32_t want) {
    constexpr auto kSpinLimit = std::chrono::milliseconds(20);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t polls = 1;; ++polls) {
      if (__atomic_load_n(h_seq, __ATOMIC_ACQUIRE) == want) return GR_OK;
      __builtin_ia32_pause();
      if ((polls & 1023u) == 0 && std::chrono::steady_clock::now() - t0 > kSpinLimit) break;
    }
    GR_HIP(hipStreamSynchronize(stream));
    if (__atomic_load_n(h_seq, __ATOMIC_ACQUIRE) != want)
      return set_error(GR_ERR_HIP, "frontier counters were not published");
    return GR_OK;
  }

  // The active list of the active bitmap, for the phases that need it.  The
  // per-block offsets in d_blk and ctr->n_active are those nextIter's scan
  // left: nothing else writes them before the list is next replaced.
  int ensure_list() {
    if (list_valid) return GR_OK;
    list_valid = true;
    if (n_active == 0) return GR_OK;
    k_compact_write<<<n_cblocks, kCompactNT, 0, stream>>>(d_act_bits, n_words, g->nV, g->off_in,
                                                          P::kHasGather ? 1 : 0, d_blk, d_cgrp,
                                                          d_active, d_escan, d_cbase);
    GR_HIP(hipGetLastError());
    return GR_OK;
  }

  int get_results() override {
    const VD* src = user_order_vd();
    GR_HIP(hipMemcpyAsync(host_vdata, src, sizeof(VD) * g->nV, hipMemcpyDeviceToHost, stream));
    GR_HIP(hipStreamSynchronize(stream));
    return GR_OK;
  }

  // d_srcval after the vertex data was written as a whole
  int refresh_source_values() {
    if constexpr (source_value_v<P>) {
      k_source_values<P><<<grid_1d(g->nV, 256) < 8192 ? grid_1d(g->nV, 256) : 8192, 256, 0,
                           stream>>>(g->nV, d_vd, d_srcval);
      GR_HIP(hipGetLastError());
      src_flags_ok = true;  // no flags
      gather_flags = false;
    }
    return GR_OK;
  }

  int reset_vertex_data() override {
    GR_TRY(drop_candidates());
    GR_HIP(hipMemcpyAsync(d_vd, d_vd_init, sizeof(VD) * g->nV, hipMemcpyDeviceToDevice, stream));
    GR_TRY(refresh_source_values());
    delta_chain = delta_ready = false;
    return GR_OK;
  }

  int set_delta_gather(int mode) override {
    delta_mode = mode;
    return GR_OK;
  }

  int read_edge_data(void* out, size_t eelem) override {
    if constexpr (kEdgeEmpty) {
      (void)out;
      return eelem == 0 ? GR_OK : set_error(GR_ERR_INVALID, "this program has no edge data");
    } else {
      if (eelem != sizeof(ED)) return set_error(GR_ERR_INVALID, "edge element size mismatch");
      const int64_t E = g->nE;
      if (E == 0) return GR_OK;
      if (streaming) {  // host copy in CSC order
        ED* o = static_cast<ED*>(out);
        for (int64_t k = 0; k < E; ++k) o[h_csc_eid[(size_t)k]] = h_ed[k];
        return GR_OK;
      }
      if constexpr (kScatter) {
        GR_HIP(hipMemcpyAsync(out, d_ed, sizeof(ED) * E, hipMemcpyDeviceToHost, stream));
      } else {
        // CSC order -> edge-list order: slot i holds edge csc_eid[i]
        ED* tmp = nullptr;
        GR_TRY(dalloc(&tmp, E));
        k_unpermute_words<<<grid_1d(E, 256), 256, 0, stream>>>(
            reinterpret_cast<const uint32_t*>(d_ed), g->csc_eid, E, (int)(sizeof(ED) / 4),
            reinterpret_cast<uint32_t*>(tmp));
        GR_HIP(hipMemcpyAsync(out, tmp, sizeof(ED) * E, hipMemcpyDeviceToHost, stream));
        GR_HIP(hipStreamSynchronize(stream));
        (void)hipFree(tmp);
      }
      GR_HIP(hipStreamSynchronize(stream));
      return GR_OK;
    }
  }

  int read_vertex_data(void* out, size_t velem) override {
    if (velem != sizeof(VD)) return set_error(GR_ERR_INVALID, "vertex element size mismatch");
    const VD* src = user_order_vd();
    GR_HIP(hipMemcpyAsync(out, src, sizeof(VD) * g->nV, hipMemcpyDeviceToHost, stream));
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
    {
      int64_t pushes = 0, scans = 0;
      sum_pairs(ev[kPhPush], &pushes, &stats.push_ms);
      sum_pairs(ev[kPhInert], &scans, &stats.inert_ms);
    }
    if (out) *out = stats;
    return GR_OK;
  }
};

}  // namespace gr

#endif
