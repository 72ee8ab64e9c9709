// Connected-components driver over the MI355X GAS engine, following
// connected_component.cu (program :25-63, loop :66-81, main :91-168).
// Boundary edits as in bfs.cpp.
#include <algorithm>
#include <climits>
#include <vector>

#include "app_common.h"
#include "graphreduce_amd/GASEngineGPU.h"

struct CC {
  typedef int VertexData;
  struct EdgeData {};
  typedef int GatherResult;
  static const int gatherZero = INT_MAX;
  static int gatherReduce(const int& left, const int& right) { return std::min(left, right); }
  static int gatherMap(const VertexData*, const VertexData* srcLabel, const EdgeData*) {
    return *srcLabel;
  }
  static bool apply(VertexData* curLabel, GatherResult label) {
    bool changed = label < *curLabel;
    *curLabel = std::min(*curLabel, label);
    return changed;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};
GR_REGISTER_PROGRAM(CC, GR_PROGRAM_CC)

using graphreduce_amd::GASEngineGPU;

template <typename Engine>
void run(int nVertices, CC::VertexData* vertexData, int nEdges, const int* srcs, const int* dsts) {
  Engine engine;
  engine.setGraph(nVertices, vertexData, nEdges, 0, srcs, dsts);
  engine.setActive(0, nVertices);
  int64_t t0 = app::currentTime();
  engine.run();
  engine.getResults();
  int64_t t1 = app::currentTime();
  printf("Took %f ms\n", (t1 - t0) / 1000.0f);
}

int main(int argc, char** argv) {
  app::Args a = app::parse(argc, argv);
  if (a.pos.size() < 1) {
    printf("Usage: cc [-t] [-d] inputfile source [outputfile]\n");
    return 1;
  }
  int nVertices;
  std::vector<int> srcs, dsts;
  app::loadGraph(a.pos[0], nVertices, srcs, dsts, nullptr);
  printf("loaded %s with %d vertices and %zd edges\n", a.pos[0].c_str(), nVertices, srcs.size());
  std::vector<int> vertexData(nVertices);
  for (int i = 0; i < nVertices; ++i) vertexData[i] = i;
  if (a.t) printf("-t: the CPU reference comparison runs in tests/test_apps.py\n");
  run<GASEngineGPU<CC> >(nVertices, vertexData.data(), (int)srcs.size(), srcs.data(), dsts.data());
  if (a.d)
    for (int i = 0; i < nVertices; ++i) printf("%d %d\n", i, vertexData[i]);
  // connected_component.cu:156 takes the output file after the input
  const size_t out_pos = a.pos.size() > 2 ? 2 : 1;
  if (a.pos.size() > out_pos) {
    printf("writing results to %s\n", a.pos[out_pos].c_str());
    FILE* f = fopen(a.pos[out_pos].c_str(), "w");
    for (int i = 0; i < nVertices; ++i) fprintf(f, "%d %d\n", i, vertexData[i]);
    fclose(f);
  }
  return 0;
}
