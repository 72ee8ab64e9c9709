// Device instantiations of the reference's four vertex programs.
//
// Each struct keeps the reference's members and semantics (names, types,
// gatherZero, gatherMap/gatherReduce/apply/scatter); the engine kernels are
// templates over them.  Three additions are engine traits, not behaviour:
//   kHasGather        false when gatherMap/gatherReduce are constants, so the
//                     engine may skip the gather (observationally identical:
//                     apply ignores the value).
//   kHasScatter       false when scatter is a no-op.
//   kGatherUsesDst    false when gatherMap ignores its dst argument.
//   kDeltaGather      true when gatherReduce is min, gatherMap reads only the
//                     source value and the edge, and apply keeps
//                     min(current, g) and reports exactly a decrease.  After an
//                     all-active iteration the engine may then gather only
//                     from sources that changed in the previous iteration:
//                     every other source was already gathered by the
//                     destination after its last change and cannot lower it,
//                     so apply sees the same outcome and value.
// and apply() receives the engine's iteration number as a third argument in
// place of the reference's __constant__ g_iterationCountGPU (bfs.cu:27-28).
#ifndef GRAPHREDUCE_AMD_GR_PROGRAMS_H
#define GRAPHREDUCE_AMD_GR_PROGRAMS_H

#include <hip/hip_runtime.h>
#include <limits.h>

namespace gr {

// bfs.cu:31-80
struct BFSProgram {
  struct VertexData {
    int depth;
  };
  struct EdgeData {};
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX - 1;
  static constexpr bool kHasGather = false;
  static constexpr bool kHasScatter = false;
  static constexpr bool kGatherUsesDst = false;
  static constexpr bool kDeltaGather = false;

  __device__ static int gatherReduce(const int&, const int&) { return 0; }
  __device__ static int gatherMap(const VertexData*, const VertexData*, const EdgeData*) {
    return 0;
  }
  // bfs.cu:59-71: first visit records the current level.
  __device__ static bool apply(VertexData* vert, int, int iteration) {
    if (vert->depth == -1) {
      vert->depth = iteration;
      return true;
    }
    return false;
  }
  __device__ static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// sssp.cu:25-66
struct SSSPProgram {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static constexpr int maxLength = 100000;
  static constexpr int gatherZero = INT_MAX - maxLength;
  static constexpr bool kHasGather = true;
  static constexpr bool kHasScatter = false;
  static constexpr bool kGatherUsesDst = false;
  static constexpr bool kDeltaGather = true;
  // An unreached source (distance gatherZero, the sssp.cu:45-52 initial
  // value) maps to gatherZero + w on an edge of weight w in [0, maxLength]:
  // at least gatherZero and without overflow, so min() with gatherZero
  // absorbs it and it can lower no distance.
  static constexpr bool kInertSources = true;
  __device__ static bool inertSource(const VertexData& d) { return d == gatherZero; }
  __device__ static bool edgeKeepsInert(const EdgeData& w) { return w >= 0 && w <= maxLength; }

  __device__ static int gatherReduce(const int& left, const int& right) {
    return min(left, right);
  }
  __device__ static int gatherMap(const VertexData*, const VertexData* srcDist,
                                  const EdgeData* edgeLen) {
    return *srcDist + *edgeLen;
  }
  __device__ static bool apply(VertexData* curDist, GatherResult dist, int) {
    bool changed = dist < *curDist;
    *curDist = min(*curDist, dist);
    return changed;
  }
  __device__ static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// pagerank.cu:27-73.  The arithmetic is written with explicit round-to-
// nearest operations so that no multiply-add is contracted into an FMA: the
// reference computes pageConst + (1 - pageConst) * g with two roundings.
struct PageRankProgram {
  static constexpr float pageConst = 0.15f;
  static constexpr float tol = 0.01f;
  struct VertexData {
    float rank;
    int numOutEdges;
  };
  struct EdgeData {};
  typedef float GatherResult;
  static constexpr float gatherZero = 0.0f;
  static constexpr bool kHasGather = true;
  static constexpr bool kHasScatter = false;
  static constexpr bool kGatherUsesDst = false;
  static constexpr bool kDeltaGather = false;

  __device__ static float gatherMap(const VertexData*, const VertexData* src, const EdgeData*) {
    return __fdiv_rn(src->rank, (float)src->numOutEdges);
  }
  // gatherMap depends on the source through rank / numOutEdges only: the
  // engine keeps that quotient per vertex (the same rounded division)
  typedef float SourceValue;
  __device__ static float sourceValue(const VertexData& v) {
    return __fdiv_rn(v.rank, (float)v.numOutEdges);
  }
  __device__ static float gatherMapValue(const float& s, const EdgeData*) { return s; }
  __device__ static float gatherReduce(const float& left, const float& right) {
    return __fadd_rn(left, right);
  }
  // the gather is a sum of non-negative terms: the blocked gather may add
  // them exactly in fixed point and round once (a deterministic order-free
  // sum; the reference's GPU sums in tree order, its CPU engine sequentially)
  static constexpr bool kFixedPointSum = true;
  __device__ static bool apply(VertexData* vertexData, const float& gatherResult, int) {
    const float damp = 1.0f - pageConst;
    float newRank = __fadd_rn(pageConst, __fmul_rn(damp, gatherResult));
    bool ret = fabsf(__fsub_rn(newRank, vertexData->rank)) >= tol;
    vertexData->rank = newRank;
    return ret;
  }
  __device__ static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// connected_component.cu:25-63
struct CCProgram {
  typedef int VertexData;
  struct EdgeData {};
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX;
  static constexpr bool kHasGather = true;
  static constexpr bool kHasScatter = false;
  static constexpr bool kGatherUsesDst = false;
  static constexpr bool kDeltaGather = true;

  __device__ static int gatherReduce(const int& left, const int& right) {
    return min(left, right);
  }
  __device__ static int gatherMap(const VertexData*, const VertexData* srcLabel, const EdgeData*) {
    return *srcLabel;
  }
  __device__ static bool apply(VertexData* curLabel, GatherResult label, int) {
    bool changed = label < *curLabel;
    *curLabel = min(*curLabel, label);
    return changed;
  }
  __device__ static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

// A program with a non-trivial scatter (the four reference apps have none),
// for the general Program::scatter of refgas.h:172-196: the SSSP gather and
// apply of sssp.cu:25-66 on links that congest.  Each time a vertex's
// distance drops, its out-edges toward vertices still farther away get one
// unit heavier.  scatter runs for every out-edge of a vertex whose apply
// returned true, after all applies of the iteration, and only writes that
// edge's data, so the calls are independent of each other.
struct SSSPCongestionProgram {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX - 100000;
  static constexpr bool kHasGather = true;
  static constexpr bool kHasScatter = true;
  static constexpr bool kGatherUsesDst = false;
  static constexpr bool kDeltaGather = false;

  __device__ static int gatherReduce(const int& left, const int& right) {
    return min(left, right);
  }
  __device__ static int gatherMap(const VertexData*, const VertexData* srcDist,
                                  const EdgeData* edgeLen) {
    return *srcDist + *edgeLen;
  }
  __device__ static bool apply(VertexData* curDist, GatherResult dist, int) {
    bool changed = dist < *curDist;
    *curDist = min(*curDist, dist);
    return changed;
  }
  __device__ static void scatter(const VertexData* src, const VertexData* dst, EdgeData* edge) {
    if (*dst > *src) *edge += 1;
  }
};

}  // namespace gr

#endif
