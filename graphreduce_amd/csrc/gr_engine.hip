// C ABI of the GAS engine (gr_abi.h) and the engine instantiations of the
// library's built-in vertex programs (gr_programs.h); the kernels and the
// Engine<Program> template are in gr_engine_impl.h.
#include "gr_engine_impl.h"
#include "gr_programs.h"

struct gr_engine {
  int program = 0;
  gr::EngineBase* impl = nullptr;
};

using namespace gr;

namespace {

EngineBase* make_engine(int program) {
  switch (program) {
    case GR_PROGRAM_BFS: return new Engine<BFSProgram>();
    case GR_PROGRAM_SSSP: return new Engine<SSSPProgram>();
    case GR_PROGRAM_PAGERANK: return new Engine<PageRankProgram>();
    case GR_PROGRAM_CC: return new Engine<CCProgram>();
    case GR_PROGRAM_SSSP_CONGESTION: return new Engine<SSSPCongestionProgram>();
    default: return nullptr;
  }
}

int check(gr_engine* e, bool need_bound = true) {
  if (!e || !e->impl) return set_error(GR_ERR_INVALID, "engine is NULL");
  if (need_bound && !e->impl->bound) return set_error(GR_ERR_STATE, "engine has no graph (call setGraph)");
  return GR_OK;
}

}  // namespace

extern "C" int gr_engine_create(int program_id, gr_engine** out) {
  if (!out) return set_error(GR_ERR_INVALID, "out is NULL");
  EngineBase* impl = make_engine(program_id);
  if (!impl) return set_errorf(GR_ERR_UNSUPPORTED, "unknown program id %d", program_id);
  gr_engine* e = new gr_engine();
  e->program = program_id;
  e->impl = impl;
  *out = e;
  return GR_OK;
}

extern "C" int gr_engine_adopt(void* engine_impl, gr_engine** out) {
  if (!out || !engine_impl) return set_error(GR_ERR_INVALID, "gr_engine_adopt: NULL argument");
  gr_engine* e = new gr_engine();
  e->program = GR_PROGRAM_CUSTOM;
  e->impl = static_cast<EngineBase*>(engine_impl);
  *out = e;
  return GR_OK;
}

extern "C" int gr_engine_destroy(gr_engine* e) {
  if (!e) return GR_OK;
  delete e->impl;
  delete e;
  return GR_OK;
}

extern "C" int gr_engine_set_graph(gr_engine* e, int32_t n_vertices, void* vertex_data,
                                   size_t vertex_elem_size, int32_t n_edges,
                                   const void* edge_data, size_t edge_elem_size,
                                   const int32_t* srcs, const int32_t* dsts) {
  GR_TRY(check(e, false));
  if (vertex_elem_size != e->impl->vertex_size())
    return set_errorf(GR_ERR_INVALID, "vertex element size %zu != program VertexData size %zu",
                      vertex_elem_size, e->impl->vertex_size());
  if (e->impl->edge_size() && (!edge_data || edge_elem_size != e->impl->edge_size()))
    return set_errorf(GR_ERR_INVALID, "program needs edge data of %zu-byte elements",
                      e->impl->edge_size());
  if (!vertex_data) return set_error(GR_ERR_INVALID, "vertex data is NULL");
  if (e->impl->stream_max_edges > 0)
    return e->impl->bind_streaming(n_vertices, vertex_data, vertex_elem_size, n_edges, edge_data,
                                   edge_elem_size, srcs, dsts);
  gr_graph* g = nullptr;
  GR_TRY(gr_graph_from_edges(n_vertices, n_edges, srcs, dsts, &g));
  // bind() takes ownership of g once its argument checks (repeated above) pass
  return e->impl->bind(g, true, vertex_data, vertex_elem_size, edge_data, edge_elem_size, false);
}

extern "C" int gr_engine_bind_graph(gr_engine* e, gr_graph* g, void* vertex_data,
                                    size_t vertex_elem_size, const void* edge_data,
                                    size_t edge_elem_size, int edge_data_on_device) {
  GR_TRY(check(e, false));
  if (!g) return set_error(GR_ERR_INVALID, "graph is NULL");
  return e->impl->bind(g, false, vertex_data, vertex_elem_size, edge_data, edge_elem_size,
                       edge_data_on_device != 0);
}

extern "C" int gr_engine_set_active(gr_engine* e, int32_t start, int32_t end) {
  GR_TRY(check(e));
  return e->impl->set_active(start, end);
}

extern "C" int gr_engine_count_active(gr_engine* e, int32_t* n_active) {
  GR_TRY(check(e));
  if (!n_active) return set_error(GR_ERR_INVALID, "n_active is NULL");
  *n_active = e->impl->n_active;
  return GR_OK;
}

extern "C" int gr_engine_gather(gr_engine* e, int) {
  GR_TRY(check(e));
  return e->impl->gather();
}

extern "C" int gr_engine_apply(gr_engine* e) {
  GR_TRY(check(e));
  return e->impl->apply();
}

extern "C" int gr_engine_scatter_activate(gr_engine* e, int have_scatter) {
  GR_TRY(check(e));
  return e->impl->scatter_activate(have_scatter != 0);
}

extern "C" int gr_engine_next_iter(gr_engine* e, int32_t* n_active) {
  GR_TRY(check(e));
  GR_TRY(e->impl->next_iter());
  if (n_active) *n_active = e->impl->n_active;
  return GR_OK;
}

extern "C" int gr_engine_get_results(gr_engine* e) {
  GR_TRY(check(e));
  return e->impl->get_results();
}

extern "C" int gr_engine_run(gr_engine* e, int32_t* iterations) {
  GR_TRY(check(e));
  int32_t it = 0;
  int rc = GR_OK;
  e->impl->run_mode = true;
  while (e->impl->n_active) {
    if ((rc = e->impl->gather()) != GR_OK) break;
    if ((rc = e->impl->apply()) != GR_OK) break;
    if ((rc = e->impl->scatter_activate(true)) != GR_OK) break;
    if ((rc = e->impl->next_iter()) != GR_OK) break;
    ++it;
  }
  e->impl->run_mode = false;
  // an apply queued by the last nextIter ran over an empty active set
  const int rc_end = e->impl->end_run();
  GR_TRY(rc);
  GR_TRY(rc_end);
  if (iterations) *iterations = it;
  return GR_OK;
}

extern "C" int gr_engine_set_iteration_count(gr_engine* e, int32_t iteration) {
  GR_TRY(check(e, false));
  e->impl->iteration = iteration;
  return GR_OK;
}

extern "C" int gr_engine_get_iteration_count(gr_engine* e, int32_t* iteration) {
  GR_TRY(check(e, false));
  if (!iteration) return set_error(GR_ERR_INVALID, "iteration is NULL");
  *iteration = e->impl->iteration;
  return GR_OK;
}

extern "C" int gr_engine_read_vertex_data(gr_engine* e, void* out, size_t vertex_elem_size) {
  GR_TRY(check(e));
  if (!out) return set_error(GR_ERR_INVALID, "out is NULL");
  return e->impl->read_vertex_data(out, vertex_elem_size);
}

extern "C" int gr_engine_read_edge_data(gr_engine* e, void* out, size_t edge_elem_size) {
  GR_TRY(check(e));
  if (!out && edge_elem_size) return set_error(GR_ERR_INVALID, "out is NULL");
  return e->impl->read_edge_data(out, edge_elem_size);
}

extern "C" int gr_program_has_scatter(int program_id) {
  switch (program_id) {
    case GR_PROGRAM_BFS: return BFSProgram::kHasScatter ? 1 : 0;
    case GR_PROGRAM_SSSP: return SSSPProgram::kHasScatter ? 1 : 0;
    case GR_PROGRAM_PAGERANK: return PageRankProgram::kHasScatter ? 1 : 0;
    case GR_PROGRAM_CC: return CCProgram::kHasScatter ? 1 : 0;
    case GR_PROGRAM_SSSP_CONGESTION: return SSSPCongestionProgram::kHasScatter ? 1 : 0;
    default: return -1;
  }
}

extern "C" int gr_engine_reset_vertex_data(gr_engine* e) {
  GR_TRY(check(e));
  return e->impl->reset_vertex_data();
}

extern "C" int gr_engine_set_direction(gr_engine* e, int mode, double alpha) {
  GR_TRY(check(e, false));
  if (mode < 0 || mode > 2) return set_error(GR_ERR_INVALID, "direction mode must be 0, 1 or 2");
  e->impl->dir_mode = mode;
  if (alpha > 0) e->impl->dir_alpha = alpha;
  return GR_OK;
}

extern "C" int gr_engine_set_push_share(gr_engine* e, double share) {
  GR_TRY(check(e, false));
  if (!(share >= 0.0)) return set_error(GR_ERR_INVALID, "push share must be >= 0");
  e->impl->push_share = share;
  return GR_OK;
}

extern "C" int gr_engine_set_blocked_gather(gr_engine* e, double min_share) {
  GR_TRY(check(e, false));
  if (!(min_share == min_share)) return set_error(GR_ERR_INVALID, "blocked-gather share is NaN");
  e->impl->blocked_share = min_share;
  return GR_OK;
}

extern "C" int gr_engine_set_fuse_share(gr_engine* e, double share) {
  GR_TRY(check(e, false));
  if (!(share >= 0.0)) return set_error(GR_ERR_INVALID, "fuse share must be >= 0");
  e->impl->fuse_share = share;
  return GR_OK;
}

extern "C" int gr_engine_set_streaming(gr_engine* e, int64_t max_edges_per_shard) {
  GR_TRY(check(e, false));
  if (max_edges_per_shard < 0) return set_error(GR_ERR_INVALID, "max_edges_per_shard < 0");
  e->impl->stream_max_edges = max_edges_per_shard;
  return GR_OK;
}

extern "C" int gr_engine_set_vertex_order(gr_engine* e, int order) {
  GR_TRY(check(e, false));
  if (order < 0 || order > 2)
    return set_error(GR_ERR_INVALID,
                     "vertex order must be 0 (as given), 1 (by degree) or 2 (by degree when "
                     "the degrees are skewed)");
  e->impl->vertex_order = order;
  return GR_OK;
}

extern "C" int gr_engine_set_delta_gather(gr_engine* e, int mode) {
  GR_TRY(check(e, false));
  if (mode < 0 || mode > 3)
    return set_error(GR_ERR_INVALID,
                     "delta gather mode must be 0 (off), 1 (adaptive), 2 (pull) or 3 (push)");
  return e->impl->set_delta_gather(mode);
}

extern "C" int gr_engine_set_gather_tile(gr_engine* e, int items_per_thread) {
  GR_TRY(check(e, false));
  if (items_per_thread != 4 && items_per_thread != 8 && items_per_thread != 16)
    return set_error(GR_ERR_INVALID, "gather items per thread must be 4, 8 or 16");
  e->impl->gather_vt = items_per_thread;
  return GR_OK;
}

extern "C" int gr_engine_enable_timing(gr_engine* e, int on) {
  GR_TRY(check(e, false));
  e->impl->timing = on != 0;
  return GR_OK;
}

extern "C" int gr_engine_get_stats(gr_engine* e, gr_engine_stats* out) {
  GR_TRY(check(e, false));
  if (!out) return set_error(GR_ERR_INVALID, "out is NULL");
  return e->impl->sync_stats(out);
}

extern "C" int gr_engine_reset_stats(gr_engine* e) {
  GR_TRY(check(e, false));
  GR_TRY(e->impl->sync_stats(nullptr));
  e->impl->stats = gr_engine_stats{};
  return GR_OK;
}
