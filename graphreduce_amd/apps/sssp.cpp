// SSSP driver over the MI355X GAS engine, following sssp.cu (program
// sssp.cu:25-66, loop sssp.cu:69-111, main sssp.cu:121-250).  Boundary edits
// as in bfs.cpp; the reference averages 10 runs (sssp.cu:80), kept here.
#include <algorithm>
#include <climits>
#include <vector>

#include "app_common.h"
#include "graphreduce_amd/GASEngineGPU.h"

struct SSSP {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static const int maxLength = 100000;
  static const int gatherZero = INT_MAX - maxLength;
  static int gatherReduce(const int& left, const int& right) { return std::min(left, right); }
  static int gatherMap(const VertexData*, const VertexData* srcDist, const EdgeData* edgeLen) {
    return *srcDist + *edgeLen;
  }
  static bool apply(VertexData* curDist, GatherResult dist) {
    bool changed = dist < *curDist;
    *curDist = std::min(*curDist, dist);
    return changed;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};
GR_REGISTER_PROGRAM(SSSP, GR_PROGRAM_SSSP)

using graphreduce_amd::GASEngineGPU;

template <typename Engine>
float run(int srcVertex, int nVertices, SSSP::VertexData* vertexData, int nEdges,
          SSSP::EdgeData* edgeData, const int* srcs, const int* dsts) {
  Engine engine;
  float elapsed = 0;
  int iteration = 0;
  for (int itr = 0; itr < 10; ++itr) {
    for (int i = 0; i < nVertices; ++i) vertexData[i] = SSSP::gatherZero;
    vertexData[srcVertex] = 0;
    engine.setGraph(nVertices, vertexData, nEdges, edgeData, srcs, dsts);
    engine.setActive(0, nVertices);
    int64_t t0 = app::currentTime();
    while (engine.countActive()) {
      engine.gather();
      engine.apply();
      engine.scatterActivate();
      engine.nextIter();
      ++iteration;
    }
    engine.getResults();
    elapsed += (app::currentTime() - t0) / 1000.0f;
  }
  printf("number of iterations: %d\n", iteration);
  return elapsed / 10;
}

int main(int argc, char** argv) {
  app::Args a = app::parse(argc, argv);
  if (a.pos.size() < 2) {
    printf("Usage: sssp [-t] [-d] [-m] inputfile source [outputfile]\n");
    return 1;
  }
  int nVertices;
  std::vector<int> srcs, dsts, edgeData;
  app::loadGraph(a.pos[0], nVertices, srcs, dsts, &edgeData);
  if (edgeData.empty() && !srcs.empty()) {
    printf("No edge data available in input file\n");
    return 1;
  }
  int sourceVertex = atoi(a.pos[1].c_str());
  if (a.m) {
    int deg;
    sourceVertex = app::maxOutDegreeVertex(nVertices, srcs, &deg);
    printf("using vertex %d with degree %d as source\n", sourceVertex, deg);
  }
  if (a.t) printf("-t: the CPU reference comparison runs in tests/test_apps.py\n");
  std::vector<int> vertexData(nVertices);
  float elapsed = run<GASEngineGPU<SSSP> >(sourceVertex, nVertices, vertexData.data(),
                                           (int)srcs.size(), edgeData.data(), srcs.data(),
                                           dsts.data());
  std::vector<int> deg = app::outDegrees(nVertices, srcs);
  long long nodes = 0, edges = 0;
  for (int i = 0; i < nVertices; ++i)
    if (vertexData[i] < SSSP::gatherZero) {
      nodes += 1;
      edges += deg[i];
    }
  printf("nodes visited: %lld edges visited: %lld\n", nodes, edges);
  printf("elapsed: %.4f ms, MTEPS: %.4f MiEdges/s\n", elapsed, (float)edges / (elapsed * 1000));
  if (a.d)
    for (int i = 0; i < nVertices; ++i) printf("%d %d\n", i, vertexData[i]);
  if (a.pos.size() > 2) {
    printf("writing results to %s\n", a.pos[2].c_str());
    FILE* f = fopen(a.pos[2].c_str(), "w");
    for (int i = 0; i < nVertices; ++i) fprintf(f, "%d %d\n", i, vertexData[i]);
    fclose(f);
  }
  return 0;
}
