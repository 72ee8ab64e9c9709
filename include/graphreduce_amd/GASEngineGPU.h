// C++ host interface of the MI355X GAS engine: the reference's class
// GASEngineGPU<Program, Int, sortEdgesForGather> (gpugas.h:1025-1809) rebuilt
// over the C ABI in gr_abi.h, so that vertex-program drivers written against
// the reference (bfs.cu, sssp.cu, pagerank.cu, connected_component.cu) keep
// their loops unchanged.
//
// Differences a maintainer will meet:
//  - The device side of the four reference programs is compiled into
//    libgraphreduce_amd; a host Program struct is tied to that device
//    instantiation once, with GR_REGISTER_PROGRAM(BFS, GR_PROGRAM_BFS), which
//    checks at compile time that its VertexData / EdgeData sizes and
//    gatherZero are those of the device twin.  Any other program runs when
//    the file including it is compiled with hipcc and includes
//    graphreduce_amd/gas_program.hpp instead: the engine is then instantiated
//    for the program itself, as the reference's template argument is.
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

#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <type_traits>

#include "gr_abi.h"

namespace graphreduce_amd {

// Maps a host Program struct to the device program compiled into the library.
template <class Program>
struct ProgramId;

// Layout and identity of the library's device programs (gr_programs.h), for
// the compile-time checks of GR_REGISTER_PROGRAM.
template <int ID>
struct DeviceTwin;
template <>
struct DeviceTwin<GR_PROGRAM_BFS> {  // bfs.cu:31-80
  static constexpr size_t vertex_bytes = 4, edge_bytes = 0;
  static constexpr double gather_zero = INT_MAX - 1;
};
template <>
struct DeviceTwin<GR_PROGRAM_SSSP> {  // sssp.cu:25-66
  static constexpr size_t vertex_bytes = 4, edge_bytes = 4;
  static constexpr double gather_zero = INT_MAX - 100000;
};
template <>
struct DeviceTwin<GR_PROGRAM_PAGERANK> {  // pagerank.cu:27-73
  static constexpr size_t vertex_bytes = 8, edge_bytes = 0;
  static constexpr double gather_zero = 0.0;
};
template <>
struct DeviceTwin<GR_PROGRAM_CC> {  // connected_component.cu:25-63
  static constexpr size_t vertex_bytes = 4, edge_bytes = 0;
  static constexpr double gather_zero = INT_MAX;
};
template <>
struct DeviceTwin<GR_PROGRAM_SSSP_CONGESTION> {
  static constexpr size_t vertex_bytes = 4, edge_bytes = 4;
  static constexpr double gather_zero = INT_MAX - 100000;
};

template <class Program, class = void>
struct has_program_id : std::false_type {};
template <class Program>
struct has_program_id<Program, std::void_t<decltype(ProgramId<Program>::value)>>
    : std::true_type {};

// kHasScatter of a program of the caller's (default true, as gas_program.hpp)
template <class P, class = void>
struct program_has_scatter : std::true_type {};
template <class P>
struct program_has_scatter<P, std::void_t<decltype(P::kHasScatter)>>
    : std::integral_constant<bool, P::kHasScatter> {};

// Defined in gas_program.hpp (hipcc): the engine instantiated for Program.
template <class Program>
int create_engine(gr_engine** out);

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

  GASEngineGPU() {
    if constexpr (has_program_id<Program>::value) {
      check(gr_engine_create(ProgramId<Program>::value, &e_), "create engine");
    } else {
#ifdef GR_CUSTOM_PROGRAMS
      check(create_engine<Program>(&e_), "create engine");
#else
      static_assert(has_program_id<Program>::value,
                    "register the program with GR_REGISTER_PROGRAM, or include "
                    "graphreduce_amd/gas_program.hpp and compile with hipcc");
#endif
    }
  }
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
    bool scatters = false;
    if constexpr (has_program_id<Program>::value)
      scatters = gr_program_has_scatter(ProgramId<Program>::value) == 1;
    else
      scatters = program_has_scatter<Program>::value;
    edgeData_ = esz && scatters ? edgeData : nullptr;
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

#define GR_REGISTER_PROGRAM(PROGRAM, ID)                                                    \
  namespace graphreduce_amd {                                                               \
  template <>                                                                               \
  struct ProgramId<PROGRAM> {                                                               \
    static constexpr int value = ID;                                                        \
  };                                                                                        \
  static_assert(sizeof(PROGRAM::VertexData) == DeviceTwin<ID>::vertex_bytes,                \
                #PROGRAM ": VertexData differs in size from the library's device program"); \
  static_assert((std::is_empty<PROGRAM::EdgeData>::value ? 0 : sizeof(PROGRAM::EdgeData)) == \
                    DeviceTwin<ID>::edge_bytes,                                             \
                #PROGRAM ": EdgeData differs in size from the library's device program");   \
  static_assert((double)PROGRAM::gatherZero == DeviceTwin<ID>::gather_zero,                 \
                #PROGRAM ": gatherZero differs from the library's device program");         \
  }

#endif
