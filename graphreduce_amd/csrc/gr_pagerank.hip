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
#include <chrono>
#include <mutex>
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
constexpr int kPRVT = 8;  // merge-path items per thread of the gather tiles

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

// Device buffers of one precision PageRank run.
template <class P>
struct PRWork {
  typedef typename P::GatherResult G;
  float* rank = nullptr;
  float* contrib = nullptr;
  G* gsum = nullptr;
  int32_t* part = nullptr;
  Part<G>* carry = nullptr;
  Part<G>* head = nullptr;
  double* partial = nullptr;
  double* total = nullptr;
};

struct PRShape {
  int nV, w, n_tiles, n_ublocks;
  explicit PRShape(const gr_graph* g)
      : nV(g->nV),
        w((int)g->nE),
        n_tiles((int)(((int64_t)g->nV + g->nE + kNT * kPRVT - 1) / (kNT * kPRVT))),
        n_ublocks((g->nV + kPRNT - 1) / kPRNT) {}
};

// The iteration loop on `st`.  timed: device time from events (ms, gather_ms);
// otherwise the host clock around the loop gives ms and gather_ms is 0.
// Every buffer in `wk` is written before it is read, so none needs clearing.
template <class P>
int pagerank_loop(gr_graph* g, const PRWork<P>& wk, hipStream_t st, bool timed, int max_iters,
                  double tol, float* ranks, double* residuals, gr_pagerank_result* res) {
  const PRShape sh(g);
  const int tile = kNT * kPRVT;
  hipEvent_t e0 = nullptr, e1 = nullptr, g0 = nullptr, g1 = nullptr;
  struct EventGuard {
    hipEvent_t* ev[4];
    ~EventGuard() {
      for (hipEvent_t* e : ev)
        if (*e) (void)hipEventDestroy(*e);
    }
  } eguard{{&e0, &e1, &g0, &g1}};
  if (timed) {
    GR_HIP(hipEventCreate(&e0));
    GR_HIP(hipEventCreate(&e1));
    GR_HIP(hipEventCreate(&g0));
    GR_HIP(hipEventCreate(&g1));
  }
  // identity active list: escan = CSC offsets, CSC base 0 for every row
  k_merge_partition<<<grid_1d((int64_t)sh.n_tiles + 1, 256), 256, 0, st>>>(
      g->off_in, sh.nV, sh.w, sh.n_tiles, tile, wk.part);
  k_pr_init<<<sh.n_ublocks, kPRNT, 0, st>>>(sh.nV, g->off_out, wk.rank, wk.contrib);
  GR_HIP(hipGetLastError());
  GR_HIP(hipStreamSynchronize(st));
  double gather_ms = 0.0, resid = 0.0;
  int it = 0;
  const auto t0 = std::chrono::steady_clock::now();
  if (timed) GR_HIP(hipEventRecord(e0, st));
  while (it < max_iters) {
    if (timed) GR_HIP(hipEventRecord(g0, st));
    if (sh.n_tiles > 0)
      k_gather_tiles<P, kPRVT><<<sh.n_tiles, kNT, 0, st>>>(
          sh.nV, sh.w, nullptr, g->off_in, wk.part, nullptr, g->srcs, wk.contrib,
          reinterpret_cast<const typename P::EdgeData*>(wk.contrib), wk.gsum, wk.carry, wk.head,
          nullptr, nullptr, nullptr);
    if (timed) GR_HIP(hipEventRecord(g1, st));
    k_gather_fixup<P><<<grid_1d(sh.n_tiles, 256), 256, 0, st>>>(sh.n_tiles, wk.carry, wk.head,
                                                               wk.gsum);
    k_pr_update<typename P::GatherResult><<<sh.n_ublocks, kPRNT, 0, st>>>(
        sh.nV, wk.gsum, g->off_out, wk.rank, wk.contrib, wk.partial);
    k_sum_partials<<<1, 1024, 0, st>>>(wk.partial, sh.n_ublocks, wk.total);
    GR_HIP(hipGetLastError());
    GR_HIP(hipMemcpyAsync(&resid, wk.total, sizeof(double), hipMemcpyDeviceToHost, st));
    GR_HIP(hipStreamSynchronize(st));
    if (timed) {
      float gm = 0.0f;
      GR_HIP(hipEventElapsedTime(&gm, g0, g1));
      gather_ms += gm;
    }
    if (residuals) residuals[it] = resid;
    ++it;
    if (resid < tol) break;
  }
  double ms = 0.0;
  if (timed) {
    GR_HIP(hipEventRecord(e1, st));
    GR_HIP(hipEventSynchronize(e1));
    float fms = 0.0f;
    GR_HIP(hipEventElapsedTime(&fms, e0, e1));
    ms = fms;
  } else {
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (ranks)
    GR_HIP(hipMemcpyAsync(ranks, wk.rank, sizeof(float) * sh.nV, hipMemcpyDeviceToHost, st));
  GR_HIP(hipStreamSynchronize(st));
  if (res) {
    res->iterations = it;
    res->residual = resid;
    res->ms = ms;
    res->gather_ms = gather_ms;
    res->gathered_edges = (int64_t)it * g->nE;
  }
  return GR_OK;
}

template <class P>
int pagerank_precise(gr_graph* g, int max_iters, double tol, float* ranks, double* residuals,
                     gr_pagerank_result* res) {
  typedef typename P::GatherResult G;
  const PRShape sh(g);
  Buf<float> rank, contrib;
  Buf<G> gsum;
  Buf<int32_t> part;
  Buf<Part<G>> carry, head;
  Buf<double> partial, total;
  GR_TRY(rank.alloc(sh.nV));
  GR_TRY(contrib.alloc(sh.nV));
  GR_TRY(gsum.alloc(sh.nV));
  GR_TRY(part.alloc((int64_t)sh.n_tiles + 1));
  GR_TRY(carry.alloc(sh.n_tiles));
  GR_TRY(head.alloc(sh.n_tiles));
  GR_TRY(partial.alloc(sh.n_ublocks));
  GR_TRY(total.alloc(1));
  hipStream_t st = nullptr;
  GR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  PRWork<P> wk;
  wk.rank = rank.p;
  wk.contrib = contrib.p;
  wk.gsum = gsum.p;
  wk.part = part.p;
  wk.carry = carry.p;
  wk.head = head.p;
  wk.partial = partial.p;
  wk.total = total.p;
  return pagerank_loop<P>(g, wk, st, true, max_iters, tol, ranks, residuals, res);
}

// ------------------------------------------------- allocation-free variant
//
// The buffers come from `scratch`, a static device array that every GPU the
// library runs on holds (its own copy) from the first use of this file's
// kernels until the process ends.  Host code never passes the array's name
// to a kernel: the device address of each GPU's copy is looked up once with
// hipGetSymbolAddress and kept per device.

constexpr size_t kPRScratchBytes = (size_t)80 << 20;
static __device__ __align__(256) unsigned char scratch[kPRScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

static int scratch_address(unsigned char** out) {
  int dev = 0;
  GR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_errorf(GR_ERR_INVALID, "device %d beyond the %d the scratch table holds", dev,
                      kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_errorf(GR_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return GR_OK;
}

// Carves the run's buffers out of [base, base + kPRScratchBytes) at 256-byte
// boundaries; with base == nullptr only counts.  Returns the bytes needed.
template <class P>
size_t carve(unsigned char* base, const PRShape& sh, PRWork<P>* wk) {
  typedef typename P::GatherResult G;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  wk->rank = reinterpret_cast<float*>(take(sizeof(float) * (size_t)sh.nV));
  wk->contrib = reinterpret_cast<float*>(take(sizeof(float) * (size_t)sh.nV));
  wk->gsum = reinterpret_cast<G*>(take(sizeof(G) * (size_t)sh.nV));
  wk->part = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * ((size_t)sh.n_tiles + 1)));
  wk->carry = reinterpret_cast<Part<G>*>(take(sizeof(Part<G>) * (size_t)sh.n_tiles));
  wk->head = reinterpret_cast<Part<G>*>(take(sizeof(Part<G>) * (size_t)sh.n_tiles));
  wk->partial = reinterpret_cast<double*>(take(sizeof(double) * (size_t)sh.n_ublocks));
  wk->total = reinterpret_cast<double*>(take(sizeof(double)));
  return off;
}

template <class P>
int pagerank_precise_noalloc(gr_graph* g, int max_iters, double tol, float* ranks,
                             double* residuals, gr_pagerank_result* res) {
  const PRShape sh(g);
  PRWork<P> wk;
  const size_t need = carve<P>(nullptr, sh, &wk);
  if (need > kPRScratchBytes)
    return set_errorf(GR_ERR_INVALID,
                      "gr_pagerank_precise_noalloc: this graph needs %zu bytes of scratch, the "
                      "maximum is %zu (16 B per vertex + 36 B per 2048 vertices and edges)",
                      need, kPRScratchBytes);
  unsigned char* base = nullptr;
  GR_TRY(scratch_address(&base));
  carve<P>(base, sh, &wk);
  // the null stream: no stream or event is created
  return pagerank_loop<P>(g, wk, nullptr, false, max_iters, tol, ranks, residuals, res);
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

extern "C" int gr_pagerank_precise_noalloc(gr_graph* g, int max_iters, double tol,
                                           int compensated, float* ranks, double* residuals,
                                           gr_pagerank_result* result) {
  if (!g || max_iters < 0)
    return set_error(GR_ERR_INVALID, "gr_pagerank_precise_noalloc: bad arguments");
  if ((int64_t)g->nV + g->nE >= ((int64_t)1 << 31) - kNT * 16)
    return set_error(GR_ERR_INVALID, "graph too large for int32 merge-path coordinates");
  if (compensated)
    return pagerank_precise_noalloc<PRKahanProgram>(g, max_iters, tol, ranks, residuals, result);
  return pagerank_precise_noalloc<PRPlainProgram>(g, max_iters, tol, ranks, residuals, result);
}
