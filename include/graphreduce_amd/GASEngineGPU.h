// C++ host interface of the MI355X GAS engine: the reference's class
// GASEngineGPU<Program, Int, sortEdgesForGather> (gpugas.h:1025-1809) rebuilt
// over the C ABI in gr_abi.h, so that vertex-program drivers written against
// the reference (bfs.cu, sssp.cu, pagerank.cu, connected_component.cu) keep
// their loops unchanged.
//
// Differences a maintainer will meet:
//  - The device side of a Program is compiled into libgraphreduce_amd (the
//    four reference programs).  A host Program struct is tied to its device
//    instantiation once, with GR_REGISTER_PROGRAM(BFS, GR_PROGRAM_BFS).
//  - BFS's level reached the device through a __constant__ variable and
//    cudaMemcpyToSymbol (bfs.cu:27-28, 83-90); here the engine hands its
//    iteration number to Program::apply, set with setIterationCount().
//  - getResults() copies the device vertex data back into the caller's
//    array (the reference's copy is compiled out, gpugas.h:1778-1787), and
//    for a program whose scatter changes edge data also the edge data (the
//    reference's CPU engine works on the caller's arrays, refgas.h:69-72).
//  - Errors print the library's message and abort(), like the reference's
//    CHECK (gpugas.h:158-165).
#ifndef GRAPHREDUCE_AMD_GASENGINEGPU_H
#define GRAPHREDUCE_AMD_GASENGINEGPU_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <type_traits>

#include "gr_abi.h"

namespace graphreduce_amd {

// Maps a host Program struct to the device program compiled into the library.
template <class Program>
struct ProgramId;

inline void check(int rc, const char* what) {
  if (rc != GR_OK) {
    fprintf(stderr, "graphreduce_amd: %s failed (%d): %s\n", what, rc, gr_last_error());
    abort();
  }
}

template <class Program, class Int = int32_t, bool sortEdgesForGather = true>
class GASEngineGPU {
  static_assert(std::is_same<Int, int32_t>::value, "vertex ids and offsets are int32");

 public:
  typedef typename Program::VertexData VertexData;
  typedef typename Program::EdgeData EdgeData;
  typedef typename Program::GatherResult GatherResult;

  GASEngineGPU() { check(gr_engine_create(ProgramId<Program>::value, &e_), "create engine"); }
  ~GASEngineGPU() { gr_engine_destroy(e_); }
  GASEngineGPU(const GASEngineGPU&) = delete;
  GASEngineGPU& operator=(const GASEngineGPU&) = delete;

  // gpugas.h:1314-1319
  void setGraph(Int nVertices, VertexData* vertexData, Int nEdges, EdgeData* edgeData,
                const Int* edgeListSrcs, const Int* edgeListDsts) {
    const size_t esz = std::is_empty<EdgeData>::value ? 0 : sizeof(EdgeData);
    check(gr_engine_set_graph(e_, nVertices, vertexData, sizeof(VertexData), nEdges,
                              esz ? edgeData : nullptr, esz, edgeListSrcs, edgeListDsts),
          "setGraph");
    edgeData_ = esz && gr_program_has_scatter(ProgramId<Program>::value) == 1 ? edgeData : nullptr;
  }

  // gpugas.h:1524
  void setActive(Int vertexStart, Int vertexEnd) {
    check(gr_engine_set_active(e_, vertexStart, vertexEnd), "setActive");
  }

  // gpugas.h:1548
  Int countActive() {
    int32_t n = 0;
    check(gr_engine_count_active(e_, &n), "countActive");
    return n;
  }

  // gpugas.h:1553
  void gather(bool haveGather = true) { check(gr_engine_gather(e_, haveGather), "gather"); }

  // gpugas.h:1621
  void apply() { check(gr_engine_apply(e_), "apply"); }

  // gpugas.h:1650
  void scatterActivate(bool haveScatter = true) {
    check(gr_engine_scatter_activate(e_, haveScatter), "scatterActivate");
  }

  // gpugas.h:1761
  Int nextIter() {
    int32_t n = 0;
    check(gr_engine_next_iter(e_, &n), "nextIter");
    return n;
  }

  // gpugas.h:1778
  void getResults() {
    check(gr_engine_get_results(e_), "getResults");
    if (edgeData_) check(gr_engine_read_edge_data(e_, edgeData_, sizeof(EdgeData)), "getResults");
  }

  // gpugas.h:1792
  void run() {
    int32_t it = 0;
    check(gr_engine_run(e_, &it), "run");
    printf("Iterations: %d\n", it);
  }

  // Level handed to Program::apply (replaces setIterationCount<true>).
  void setIterationCount(Int iteration) {
    check(gr_engine_set_iteration_count(e_, iteration), "setIterationCount");
  }

  // push/pull switch: 0 automatic, 1 push only, 2 pull only
  void setDirection(int mode, double alpha = 15.0) {
    check(gr_engine_set_direction(e_, mode, alpha), "setDirection");
  }

  gr_engine* handle() { return e_; }

 private:
  gr_engine* e_ = nullptr;
  EdgeData* edgeData_ = nullptr;  // written back by getResults (programs with a scatter)
};

}  // namespace graphreduce_amd

#define GR_REGISTER_PROGRAM(PROGRAM, ID)                    \
  namespace graphreduce_amd {                               \
  template <>                                               \
  struct ProgramId<PROGRAM> {                               \
    static constexpr int value = ID;                        \
  };                                                        \
  }

#endif
