// GASEngineGPU.h with a program whose scatter changes edge data: the loop of
// sssp.cu over the congestion program; getResults() must write the edge data
// back into the caller's array.  Prints "dist" and "w" lines for the test.
#include <stdio.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "graphreduce_amd/GASEngineGPU.h"

struct Congestion {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static const int gatherZero = INT_MAX - 100000;
  static int gatherReduce(const int& a, const int& b) { return std::min(a, b); }
  static int gatherMap(const VertexData*, const VertexData* s, const EdgeData* w) { return *s + *w; }
  static bool apply(VertexData* cur, GatherResult d) {
    bool changed = d < *cur;
    *cur = std::min(*cur, d);
    return changed;
  }
  static void scatter(const VertexData* s, const VertexData* d, EdgeData* w) {
    if (*d > *s) *w += 1;
  }
};
GR_REGISTER_PROGRAM(Congestion, GR_PROGRAM_SSSP_CONGESTION)

int main() {
  const int n = 4;
  std::vector<int> srcs = {0, 0, 1, 2}, dsts = {1, 2, 3, 3}, w = {1, 5, 1, 1};
  std::vector<int> dist(n, Congestion::gatherZero);
  dist[0] = 0;
  graphreduce_amd::GASEngineGPU<Congestion> engine;
  engine.setGraph(n, dist.data(), (int)srcs.size(), w.data(), srcs.data(), dsts.data());
  engine.setActive(0, n);
  int trips = 0;
  while (engine.countActive()) {
    engine.gather();
    engine.apply();
    engine.scatterActivate();
    engine.nextIter();
    ++trips;
  }
  engine.getResults();
  printf("trips %d\ndist", trips);
  for (int x : dist) printf(" %d", x);
  printf("\nw");
  for (int x : w) printf(" %d", x);
  printf("\n");
  return 0;
}
