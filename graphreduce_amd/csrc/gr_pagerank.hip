// PageRank precision form on the merge-path gather.
//
// The reference's PageRank (pagerank.cu:27-101) activates a vertex's
// out-neighbours when its rank moved by at least tol = 0.01, so which
// vertices run in later iterations depends on FP32 rounding.  The precision
// form runs the same recurrence with every vertex active:
//   r_0[v] = 0.15,  r_{t+1}[v] = 0.15 + 0.85 * sum_{u->v} r_t[u] / outdeg[u]
// (pagerank.cu:46-66 arithmetic: FP32 division per source, FP32 multiply and
// add with two roundings).  The in-edge sum is either a plain FP32 tree sum
// or a compensated sum (TwoSum per addition, Knuth TAOCP vol. 2 4.2.2; the
// (sum, error) pairs of different lanes and tiles are joined the same way).
// The L1 residual sum_v |r_{t+1}[v] - r_t[v]| is accumulated in FP64 with a
// fixed reduction order and decides convergence (stop when it drops below
// tol or after max_iters steps).
#include <vector>

#include "gr_device.h"
#include "gr_gather.h"

namespace gr {

struct KahanF {
  float s;  // running sum
  float c;  // accumulated rounding error of the additions into s
};

// Error-free transformation: a + b = s + e exactly (no FMA contraction: the
// library is built with -ffp-contract=off).
__device__ __forceinline__ void two_sum(float a, float b, float* s, float* e) {
  const float x = a + b;
  const float bb = x - a;
  *e = (a - (x - bb)) + (b - bb);
  *s = x;
}

// Vertex data of the gather: the source's contribution r[u] / outdeg[u].
struct PRKahanProgram {
  typedef float VertexData;
  struct EdgeData {};
  typedef KahanF GatherResult;
  static constexpr KahanF gatherZero = {0.0f, 0.0f};
  static constexpr bool kHasGather = true;
  static constexpr bool kHasScatter = false;
  static constexpr bool kGatherUsesDst = false;
  __device__ static KahanF gatherMap(const float*, const float* src, const EdgeData*) {
    return KahanF{*src, 0.0f};
  }
  __device__ static KahanF gatherReduce(const KahanF& a, const KahanF& b) {
    KahanF r;
    float e;
    two_sum(a.s, b.s, &r.s, &e);
    r.c = (a.c + b.c) + e;
    return r;
  }
};

struct PRPlainProgram {
  typedef float VertexData;
  struct EdgeData {};
  typedef float GatherResult;
  static constexpr float gatherZero = 0.0f;
  static constexpr bool kHasGather = true;
  static constexpr bool kHasScatter = false;
  static constexpr bool kGatherUsesDst = false;
  __device__ static float gatherMap(const float*, const float* src, const EdgeData*) {
    return *src;
  }
  __device__ static float gatherReduce(const float& a, const float& b) { return a + b; }
};

__device__ __forceinline__ float gather_value(const KahanF& g) { return g.s + g.c; }
__device__ __forceinline__ float gather_value(float g) { return g; }

constexpr int kPRNT = 256;

__global__ void k_pr_init(int nV, const int32_t* __restrict__ off_out, float* __restrict__ rank,
                          float* __restrict__ contrib) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const float r = 0.15f;
  const int d = off_out[v + 1] - off_out[v];
  rank[v] = r;
  contrib[v] = d ? __fdiv_rn(r, (float)d) : 0.0f;
}

// r' = 0.15 + 0.85 * g; per-block FP64 partial of |r' - r| (fixed order).
template <class G>
__global__ __launch_bounds__(kPRNT) void k_pr_update(int nV, const G* __restrict__ gsum,
                                                     const int32_t* __restrict__ off_out,
                                                     float* __restrict__ rank,
                                                     float* __restrict__ contrib,
                                                     double* __restrict__ partial) {
  __shared__ double s_w[kPRNT / kWave];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double diff = 0.0;
  if (v < nV) {
    const float g = gather_value(gsum[v]);
    const float damp = 1.0f - 0.15f;
    const float nr = __fadd_rn(0.15f, __fmul_rn(damp, g));
    const float old = rank[v];
    diff = fabs((double)nr - (double)old);
    rank[v] = nr;
    const int d = off_out[v + 1] - off_out[v];
    contrib[v] = d ? __fdiv_rn(nr, (float)d) : 0.0f;
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) diff += __shfl_xor(diff, off, kWave);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = diff;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kPRNT / kWave; ++i) t += s_w[i];
    partial[blockIdx.x] = t;
  }
}

// Single workgroup: sum of the block partials in a fixed order.
__global__ __launch_bounds__(1024) void k_sum_partials(const double* __restrict__ partial, int n,
                                                       double* __restrict__ out) {
  __shared__ double s_w[1024 / kWave];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) t += partial[i];
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, kWave);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < 1024 / kWave; ++i) s += s_w[i];
    *out = s;
  }
}

template <class T>
struct Buf {
  T* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  int alloc(int64_t n) {
    if (n < 1) n = 1;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
      p = nullptr;
      return set_errorf(GR_ERR_NOMEM, "hipMalloc(%lld bytes): %s", (long long)(sizeof(T) * n),
                        hipGetErrorString(e));
    }
    return GR_OK;
  }
};

template <class P>
int pagerank_precise(gr_graph* g, int max_iters, double tol, float* ranks, double* residuals,
                     gr_pagerank_result* res) {
  typedef typename P::GatherResult G;
  const int nV = g->nV;
  const int w = (int)g->nE;
  constexpr int VT = 8;
  const int tile = kNT * VT;
  const int n_tiles = (int)(((int64_t)nV + w + tile - 1) / tile);
  const int n_ublocks = (nV + kPRNT - 1) / kPRNT;
  Buf<float> rank, contrib;
  Buf<G> gsum;
  Buf<int32_t> cbase, part;
  Buf<Part<G>> carry, head;
  Buf<double> partial, total;
  GR_TRY(rank.alloc(nV));
  GR_TRY(contrib.alloc(nV));
  GR_TRY(gsum.alloc(nV));
  GR_TRY(cbase.alloc((int64_t)nV + 1));
  GR_TRY(part.alloc((int64_t)n_tiles + 1));
  GR_TRY(carry.alloc(n_tiles));
  GR_TRY(head.alloc(n_tiles));
  GR_TRY(partial.alloc(n_ublocks));
  GR_TRY(total.alloc(1));
  hipStream_t st = nullptr;
  GR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  hipEvent_t e0, e1, g0, g1;
  GR_HIP(hipEventCreate(&e0));
  GR_HIP(hipEventCreate(&e1));
  GR_HIP(hipEventCreate(&g0));
  GR_HIP(hipEventCreate(&g1));
  // identity active list: escan = CSC offsets, CSC base 0 for every row
  GR_HIP(hipMemsetAsync(cbase.p, 0, sizeof(int32_t) * ((size_t)nV + 1), st));
  k_merge_partition<<<grid_1d((int64_t)n_tiles + 1, 256), 256, 0, st>>>(g->off_in, nV, w, n_tiles,
                                                                       tile, part.p);
  k_pr_init<<<n_ublocks, kPRNT, 0, st>>>(nV, g->off_out, rank.p, contrib.p);
  GR_HIP(hipGetLastError());
  GR_HIP(hipStreamSynchronize(st));
  double gather_ms = 0.0, resid = 0.0;
  int it = 0;
  GR_HIP(hipEventRecord(e0, st));
  while (it < max_iters) {
    GR_HIP(hipEventRecord(g0, st));
    if (n_tiles > 0)
      k_gather_tiles<P, VT><<<n_tiles, kNT, 0, st>>>(nV, w, nullptr, g->off_in, part.p, cbase.p,
                                                     g->srcs, contrib.p,
                                                     reinterpret_cast<const typename P::EdgeData*>(contrib.p),
                                                     gsum.p, carry.p, head.p, nullptr, nullptr,
                                                     nullptr);
    GR_HIP(hipEventRecord(g1, st));
    k_gather_fixup<P><<<grid_1d(n_tiles, 256), 256, 0, st>>>(n_tiles, carry.p, head.p, gsum.p);
    k_pr_update<G><<<n_ublocks, kPRNT, 0, st>>>(nV, gsum.p, g->off_out, rank.p, contrib.p,
                                                partial.p);
    k_sum_partials<<<1, 1024, 0, st>>>(partial.p, n_ublocks, total.p);
    GR_HIP(hipGetLastError());
    GR_HIP(hipMemcpyAsync(&resid, total.p, sizeof(double), hipMemcpyDeviceToHost, st));
    GR_HIP(hipStreamSynchronize(st));
    float gm = 0.0f;
    GR_HIP(hipEventElapsedTime(&gm, g0, g1));
    gather_ms += gm;
    if (residuals) residuals[it] = resid;
    ++it;
    if (resid < tol) break;
  }
  GR_HIP(hipEventRecord(e1, st));
  GR_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  GR_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(g0);
  (void)hipEventDestroy(g1);
  if (ranks) GR_HIP(hipMemcpy(ranks, rank.p, sizeof(float) * nV, hipMemcpyDeviceToHost));
  if (res) {
    res->iterations = it;
    res->residual = resid;
    res->ms = ms;
    res->gather_ms = gather_ms;
    res->gathered_edges = (int64_t)it * g->nE;
  }
  return GR_OK;
}

}  // namespace gr

using namespace gr;

extern "C" int gr_pagerank_precise(gr_graph* g, int max_iters, double tol, int compensated,
                                   float* ranks, double* residuals, gr_pagerank_result* result) {
  if (!g || max_iters < 0) return set_error(GR_ERR_INVALID, "gr_pagerank_precise: bad arguments");
  if ((int64_t)g->nV + g->nE >= ((int64_t)1 << 31) - kNT * 16)
    return set_error(GR_ERR_INVALID, "graph too large for int32 merge-path coordinates");
  if (compensated) return pagerank_precise<PRKahanProgram>(g, max_iters, tol, ranks, residuals, result);
  return pagerank_precise<PRPlainProgram>(g, max_iters, tol, ranks, residuals, result);
}
