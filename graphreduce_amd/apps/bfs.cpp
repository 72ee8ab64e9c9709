// BFS driver over the MI355X GAS engine, following bfs.cu (vertex program
// bfs.cu:31-80, loop bfs.cu:93-134, main bfs.cu:144-275).
//
// Edits against the reference, all at the boundary: the Program struct
// loses its __host__ __device__ qualifiers (its device twin is compiled into
// libgraphreduce_amd and tied to it by GR_REGISTER_PROGRAM); the level goes
// to the engine with setIterationCount() instead of cudaMemcpyToSymbol; -t
// (compare with the CPU reference engine) is served by the test suite
// (tests/test_apps.py runs this binary against the CPU oracle).
#include <climits>
#include <vector>

#include "app_common.h"
#include "graphreduce_amd/GASEngineGPU.h"

int g_iterationCount;

struct BFS {
  struct VertexData {
    int depth;
  };
  struct EdgeData {};
  typedef int GatherResult;
  static const int gatherZero = INT_MAX - 1;
  static int gatherReduce(const int&, const int&) { return 0; }
  static int gatherMap(const VertexData*, const VertexData*, const EdgeData*) { return 0; }
  static bool apply(VertexData* vert, int) {
    if (vert->depth == -1) {
      vert->depth = g_iterationCount;
      return true;
    }
    return false;
  }
  static void scatter(const VertexData*, const VertexData*, EdgeData*) {}
};
GR_REGISTER_PROGRAM(BFS, GR_PROGRAM_BFS)

using graphreduce_amd::GASEngineGPU;

template <typename Engine>
float run(int nVertices, BFS::VertexData* vertexData, int nEdges, const int* srcs,
          const int* dsts, int sourceVertex, int* iterations) {
  Engine engine;
  for (int i = 0; i < nVertices; ++i) vertexData[i].depth = -1;
  engine.setGraph(nVertices, vertexData, nEdges, 0, srcs, dsts);
  engine.setActive(sourceVertex, sourceVertex + 1);
  int iteration = 0;
  engine.setIterationCount(iteration);
  int64_t t0 = app::currentTime();
  while (engine.countActive()) {
    engine.gather(false);
    engine.apply();
    engine.scatterActivate(false);
    engine.nextIter();
    engine.setIterationCount(++iteration);
  }
  engine.getResults();
  int64_t t1 = app::currentTime();
  printf("search depth (number of iterations): %d\n", iteration);
  *iterations = iteration;
  return (t1 - t0) / 1000.0f;
}

int main(int argc, char** argv) {
  app::Args a = app::parse(argc, argv);
  if (a.pos.size() < 2) {
    printf("Usage: bfs [-t] [-d] [-m] inputfile source [outputFilename]\n");
    return 1;
  }
  int nVertices;
  std::vector<int> srcs, dsts;
  app::loadGraph(a.pos[0], nVertices, srcs, dsts, nullptr);
  int sourceVertex = atoi(a.pos[1].c_str());
  if (a.m) {
    int deg;
    sourceVertex = app::maxOutDegreeVertex(nVertices, srcs, &deg);
    printf("using vertex %d with degree %d as source\n", sourceVertex, deg);
  }
  if (a.t) printf("-t: the CPU reference comparison runs in tests/test_apps.py\n");
  std::vector<BFS::VertexData> vertexData(nVertices);
  int iterations = 0;
  float elapsed = run<GASEngineGPU<BFS> >(nVertices, vertexData.data(), (int)srcs.size(),
                                          srcs.data(), dsts.data(), sourceVertex, &iterations);
  std::vector<int> deg = app::outDegrees(nVertices, srcs);
  long long nodes = 0, edges = 0;
  for (int i = 0; i < nVertices; ++i)
    if (vertexData[i].depth > -1) {
      nodes += 1;
      edges += deg[i];
    }
  printf("nodes visited: %lld edges visited: %lld\n", nodes, edges);
  printf("elapsed: %.4f ms, MTEPS: %.4f MiEdges/s\n", elapsed, (float)edges / (elapsed * 1000));
  if (a.d)
    for (int i = 0; i < nVertices; ++i) printf("%d %d\n", i, vertexData[i].depth);
  if (a.pos.size() > 2) {
    printf("writing results to %s\n", a.pos[2].c_str());
    FILE* f = fopen(a.pos[2].c_str(), "w");
    for (int i = 0; i < nVertices; ++i) fprintf(f, "%d %d\n", i, vertexData[i].depth);
    fclose(f);
  }
  return 0;
}
