// Vertex programs of your own on the MI355X GAS engine.
//
// The reference takes any Program as the template argument of its engines
// (GASEngineGPU<Program>, gpugas.h:1025-1027; GASEngineRef<Program>,
// refgas.h:30-31): gatherMap / gatherReduce / apply / scatter as static
// __host__ __device__ functions plus VertexData, EdgeData, GatherResult and
// gatherZero.  Compile a file that includes this header with hipcc
// (--offload-arch=gfx950, -I<repo>/include, link -lgraphreduce_amd) and
// GASEngineGPU<YourProgram> instantiates the engine's kernels for it in your
// own code; no change to the library is needed.  A program registered with
// GR_REGISTER_PROGRAM still uses the library's built-in instantiation.
//
// Optional members of a program (defaults in brackets), the engine's traits:
//   kHasGather [true]       false: gatherMap/gatherReduce are constants and the
//                           gather may be skipped (apply ignores its value)
//   kHasScatter [true]      false: scatter is a no-op (no pass over out-edges)
//   kGatherUsesDst [true]   false: gatherMap ignores its dst argument
//   kDeltaGather [false]    true: gatherReduce is min, gatherMap reads only the
//                           source value and the edge, apply keeps
//                           min(current, g) and reports exactly a decrease
//   apply may take the iteration number as a third argument (the reference's
//   BFS reads it from a __constant__ variable, bfs.cu:27-28).
// The engine object lives in your binary; the graph, the C ABI calls and
// their errors are the library's, as for the built-in programs.
#ifndef GRAPHREDUCE_AMD_GAS_PROGRAM_HPP
#define GRAPHREDUCE_AMD_GAS_PROGRAM_HPP

#define GR_CUSTOM_PROGRAMS 1

#include <new>
#include <type_traits>
#include <utility>

#include "../../graphreduce_amd/csrc/gr_engine_impl.h"
#include "GASEngineGPU.h"
#include "gr_abi.h"

namespace graphreduce_amd {
namespace detail {

#define GR_DEFAULT_TRAIT(NAME, DEFAULT)                                            \
  template <class P, class = void>                                                \
  struct NAME##_or_default : std::integral_constant<bool, DEFAULT> {};            \
  template <class P>                                                              \
  struct NAME##_or_default<P, std::void_t<decltype(P::NAME)>>                     \
      : std::integral_constant<bool, P::NAME> {};
GR_DEFAULT_TRAIT(kHasGather, true)
GR_DEFAULT_TRAIT(kHasScatter, true)
GR_DEFAULT_TRAIT(kGatherUsesDst, true)
#undef GR_DEFAULT_TRAIT

template <class P, class = void>
struct apply_takes_iteration : std::false_type {};
template <class P>
struct apply_takes_iteration<
    P, std::void_t<decltype(P::apply(std::declval<typename P::VertexData*>(),
                                     std::declval<const typename P::GatherResult&>(), 0))>>
    : std::true_type {};

}  // namespace detail

// The program as the engine kernels see it: the user's members, the traits
// with their defaults, and apply with the iteration argument.
template <class P>
struct DeviceProgram : P {
  typedef typename P::VertexData VertexData;
  typedef typename P::EdgeData EdgeData;
  typedef typename P::GatherResult GatherResult;
  static constexpr bool kHasGather = detail::kHasGather_or_default<P>::value;
  static constexpr bool kHasScatter = detail::kHasScatter_or_default<P>::value;
  static constexpr bool kGatherUsesDst = detail::kGatherUsesDst_or_default<P>::value;
  __device__ static bool apply(VertexData* v, const GatherResult& g, int iteration) {
    if constexpr (detail::apply_takes_iteration<P>::value)
      return P::apply(v, g, iteration);
    else
      return P::apply(v, g);
  }
};

// A new engine for program P, instantiated in the calling translation unit.
template <class P>
int create_engine(gr_engine** out) {
  static_assert(std::is_trivially_copyable<typename P::VertexData>::value,
                "VertexData is copied to and from the device as bytes");
  gr::EngineBase* impl = new (std::nothrow) gr::Engine<DeviceProgram<P>>();
  if (!impl) return GR_ERR_NOMEM;
  const int rc = gr_engine_adopt(impl, out);
  if (rc != GR_OK) delete impl;
  return rc;
}

}  // namespace graphreduce_amd

#endif
