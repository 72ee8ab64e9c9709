// PageRank driver over the MI355X GAS engine, following pagerank.cu (program
// pagerank.cu:27-73, loop pagerank.cu:85-101, main pagerank.cu:104-184).
// Boundary edits as in bfs.cpp, plus: in-class `static const float`
// initialisers (pagerank.cu:29-30, 43) become constexpr, which clang needs.
#include <cmath>
#include <vector>

#include "app_common.h"
#include "graphreduce_amd/GASEngineGPU.h"

struct PageRank {
  static constexpr float pageConst = 0.15f;
  static constexpr float tol = 0.01f;
  struct VertexData {
    float rank;
    int numOutEdges;
  };
  struct EdgeData {};
  typedef float GatherResult;
  static constexpr float gatherZero = 0.0f;
  static float gatherMap(const VertexData*, const VertexData* src, const EdgeData*) {
    return src->rank / src->numOutEdges;
  }
  static float gatherReduce(const float& left, const float& right) { return left + right; }
  static bool apply(VertexData* vertexData, const float& gatherResult) {
    float newRank = pageConst + (1.0f - pageConst) * gatherResult;
    bool ret = std::fabs(newRank - vertexData->rank) >= tol;
    vertexData->rank = newRank;
    return ret;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};
GR_REGISTER_PROGRAM(PageRank, GR_PROGRAM_PAGERANK)

using graphreduce_amd::GASEngineGPU;

template <typename Engine>
void run(int nVertices, PageRank::VertexData* vertexData, int nEdges, const int* srcs,
         const int* dsts) {
  for (int i = 0; i < nVertices; ++i) vertexData[i].rank = PageRank::pageConst;
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
    printf("Usage: pagerank [-t] [-d] inputfile [outputfile]\n");
    return 1;
  }
  int nVertices;
  std::vector<int> srcs, dsts;
  app::loadGraph(a.pos[0], nVertices, srcs, dsts, nullptr);
  printf("loaded %s with %d vertices and %zd edges\n", a.pos[0].c_str(), nVertices, srcs.size());
  std::vector<int> deg = app::outDegrees(nVertices, srcs);
  std::vector<PageRank::VertexData> vertexData(nVertices);
  for (int i = 0; i < nVertices; ++i) vertexData[i].numOutEdges = deg[i];
  if (a.t) printf("-t: the CPU reference comparison runs in tests/test_apps.py\n");
  run<GASEngineGPU<PageRank> >(nVertices, vertexData.data(), (int)srcs.size(), srcs.data(),
                               dsts.data());
  if (a.d)
    for (int i = 0; i < nVertices; ++i) printf("%d %f\n", i, vertexData[i].rank);
  if (a.pos.size() > 1) {
    FILE* f = fopen(a.pos[1].c_str(), "w");
    printf("writing results to file %s\n", a.pos[1].c_str());
    for (int i = 0; i < nVertices; ++i) fprintf(f, "%d %f\n", i, vertexData[i].rank);
    fclose(f);
  }
  return 0;
}
