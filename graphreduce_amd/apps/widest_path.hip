// Widest (maximum-bottleneck) paths: a vertex program that is not one of the
// library's built-in programs, run through the engine's open Program
// interface.  The program is written in the reference's form (static
// __host__ __device__ members, two-argument apply, bfs.cu:31-80 style) and
// this file is compiled with hipcc against graphreduce_amd/gas_program.hpp,
// which instantiates the engine for it here (the reference's
// GASEngineGPU<Program>, gpugas.h:1025-1027).  Driver as sssp.cu's: all
// vertices active first, the source's width INT_MAX, the others 0;
// capacities are the edge data (hash weights on [1, 64] for synthetic graphs).
//
// Usage: widest_path [-m] inputfile source [outputfile]
#include <algorithm>
#include <climits>
#include <vector>

#include "app_common.h"
#include "graphreduce_amd/gas_program.hpp"

struct WidestPath {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static constexpr int gatherZero = 0;
  // the engine may skip the scatter pass and call gatherMap without dst
  static constexpr bool kHasScatter = false;
  static constexpr bool kGatherUsesDst = false;

  __host__ __device__ static int gatherReduce(const int& left, const int& right) {
    return left > right ? left : right;
  }
  __host__ __device__ static int gatherMap(const VertexData*, const VertexData* srcWidth,
                                           const EdgeData* capacity) {
    return *srcWidth < *capacity ? *srcWidth : *capacity;
  }
  __host__ __device__ static bool apply(VertexData* curWidth, GatherResult width) {
    if (width > *curWidth) {
      *curWidth = width;
      return true;
    }
    return false;
  }
  __host__ __device__ static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};

using graphreduce_amd::GASEngineGPU;

int main(int argc, char** argv) {
  app::Args a = app::parse(argc, argv);
  if (a.pos.size() < 2) {
    printf("Usage: widest_path [-m] inputfile source [outputfile]\n");
    return 1;
  }
  int nVertices;
  std::vector<int> srcs, dsts, caps;
  app::loadGraph(a.pos[0], nVertices, srcs, dsts, &caps);
  if (caps.empty() && !srcs.empty()) {
    printf("No edge data available in input file\n");
    return 1;
  }
  int source = atoi(a.pos[1].c_str());
  if (a.m) {
    int deg;
    source = app::maxOutDegreeVertex(nVertices, srcs, &deg);
    printf("using vertex %d with degree %d as source\n", source, deg);
  }
  std::vector<int> width(nVertices, WidestPath::gatherZero);
  width[source] = INT_MAX;
  GASEngineGPU<WidestPath> engine;
  engine.setGraph(nVertices, width.data(), (int)srcs.size(), caps.data(), srcs.data(),
                  dsts.data());
  engine.setActive(0, nVertices);
  int64_t t0 = app::currentTime();
  int iteration = 0;
  while (engine.countActive()) {
    engine.gather();
    engine.apply();
    engine.scatterActivate();
    engine.nextIter();
    ++iteration;
  }
  engine.getResults();
  const float elapsed = (app::currentTime() - t0) / 1000.0f;
  long long reached = 0;
  for (int i = 0; i < nVertices; ++i) reached += width[i] > 0;
  printf("number of iterations: %d\n", iteration);
  printf("vertices reached: %lld\n", reached);
  printf("Took: %.4f ms\n", elapsed);
  if (a.pos.size() > 2) {
    printf("writing results to %s\n", a.pos[2].c_str());
    FILE* f = fopen(a.pos[2].c_str(), "w");
    for (int i = 0; i < nVertices; ++i) fprintf(f, "%d %d\n", i, width[i]);
    fclose(f);
  }
  return 0;
}
