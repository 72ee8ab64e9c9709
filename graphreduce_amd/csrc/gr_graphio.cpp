// Graph file readers and writers (graphio.cpp:69-473) and the CSR conversion
// of util.cuh:79-108, host C++ over zlib.  Text files are read through
// gzFile, which passes uncompressed files through unchanged, as
// graphio.cpp:52-62 does; .gr files are read the same way.
#include <zlib.h>

#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/graphreduce_amd/gr_graphio.h"
#include "gr_internal.h"

struct gr_graphfile {
  int32_t n_vertices = 0;
  std::vector<int32_t> srcs, dsts, values;
  bool has_values = false;
};

namespace {

using gr::set_error;
using gr::set_errorf;

enum Symmetry { kGeneral, kSymmetric, kSkewSymmetric, kHermitian };

struct GzFile {
  gzFile f = nullptr;
  ~GzFile() {
    if (f) gzclose(f);
  }
};

bool blank_line(const std::string& s) {
  for (char c : s)
    if (!std::isspace((unsigned char)c)) return false;
  return true;
}

// One line of any length (without its end-of-line); false at end of file.
bool read_line(gzFile f, std::string& line, bool* failed) {
  line.clear();
  char buf[4096];
  for (;;) {
    if (!gzgets(f, buf, sizeof buf)) {
      int err = Z_OK;
      gzerror(f, &err);
      if (err != Z_OK && err != Z_STREAM_END) *failed = true;
      return !line.empty();
    }
    line += buf;
    if (!line.empty() && line.back() == '\n') {
      line.pop_back();
      return true;
    }
  }
}

// Integer field as sscanf("%d") reads it; false if there is none or it does
// not fit an int.
bool parse_int(const char*& p, long long* out) {
  char* end = nullptr;
  errno = 0;
  const long long v = std::strtoll(p, &end, 10);
  if (end == p || errno == ERANGE || v < INT_MIN || v > INT_MAX) return false;
  *out = v;
  p = end;
  return true;
}

// Edge value as graphio.cpp:140-146 reads it: "%f" into a float, stored into
// an int vector (truncation toward zero).
bool parse_value(const char* p, bool negate, int32_t* out) {
  char* end = nullptr;
  float f = std::strtof(p, &end);
  if (end == p || !std::isfinite(f)) return false;
  if (negate) f = -f;
  if (!(f > -2147483649.0f && f < 2147483648.0f)) return false;
  *out = (int32_t)f;
  return true;
}

// graphio.cpp:69-166: line-based edge lists.  first_line is the file's line
// number of the next line (after a header already consumed).
int load_text(gzFile f, const char* path, int first_line, char comment, bool skip_size_line,
              bool one_based, Symmetry sym, bool want_values, gr_graphfile* g) {
  std::string line;
  bool failed = false, size_line = skip_size_line;
  long long max_id = 0;
  for (long long line_no = first_line; read_line(f, line, &failed); ++line_no) {
    if (line.empty() || line[0] == comment || blank_line(line)) continue;
    if (size_line) {  // "rows cols entries": not used (graphio.cpp:110-116)
      size_line = false;
      continue;
    }
    const char* p = line.c_str();
    long long s, d;
    if (!parse_int(p, &s) || !parse_int(p, &d))
      return set_errorf(GR_ERR_INVALID, "%s:%lld: error parsing src/dst", path, line_no);
    if (one_based) {
      --s;
      --d;
    }
    if (s < 0 || d < 0 || s >= INT_MAX || d >= INT_MAX)
      return set_errorf(GR_ERR_INVALID, "%s:%lld: vertex id out of range", path, line_no);
    if (s == d) {
      // self-loop: dropped; with values requested the line does not count
      // toward the vertex count either (graphio.cpp:129-137)
      if (want_values) continue;
    } else {
      g->srcs.push_back((int32_t)s);
      g->dsts.push_back((int32_t)d);
      if (sym != kGeneral) {
        g->srcs.push_back((int32_t)d);
        g->dsts.push_back((int32_t)s);
      }
      if (want_values) {
        int32_t v, rv = 0;
        if (!parse_value(p, false, &v) ||
            (sym == kSkewSymmetric && !parse_value(p, true, &rv)))
          return set_errorf(GR_ERR_INVALID, "%s:%lld: missing or invalid edge value", path,
                            line_no);
        g->values.push_back(v);
        if (sym == kSymmetric) g->values.push_back(v);
        if (sym == kSkewSymmetric) g->values.push_back(rv);
      }
    }
    if (s > max_id) max_id = s;
    if (d > max_id) max_id = d;
  }
  if (failed) return set_errorf(GR_ERR_INVALID, "%s: read error (corrupt or truncated .gz?)", path);
  if (g->srcs.size() > (size_t)INT_MAX)
    return set_errorf(GR_ERR_INVALID, "%s: more than 2^31 - 1 edges", path);
  g->n_vertices = (int32_t)(max_id + 1);  // graphio.cpp:161-165
  g->has_values = want_values;
  return GR_OK;
}

// graphio.cpp:184-261: MatrixMarket header, then the common reader.
int load_mtx(gzFile f, const char* path, bool want_values, gr_graphfile* g) {
  std::string header;
  bool failed = false;
  if (!read_line(f, header, &failed) || failed)
    return set_errorf(GR_ERR_INVALID, "%s: error reading header", path);
  std::vector<std::string> tok;
  {
    size_t i = 0;
    while (i < header.size()) {
      while (i < header.size() && std::strchr(" \t\r", header[i])) ++i;
      size_t j = i;
      while (j < header.size() && !std::strchr(" \t\r", header[j])) ++j;
      if (j > i) tok.push_back(header.substr(i, j - i));
      i = j;
    }
  }
  // the first token ("%%MatrixMarket") is skipped unread, as graphio.cpp:207
  if (tok.size() < 5) return set_errorf(GR_ERR_INVALID, "%s: truncated MatrixMarket header", path);
  if (tok[1] != "matrix")
    return set_errorf(GR_ERR_INVALID, "%s: unrecognized token %s in header", path, tok[1].c_str());
  if (tok[2] != "coordinate")
    return set_errorf(GR_ERR_INVALID, "%s: only coordinate format is supported", path);
  if (tok[3] == "pattern")
    want_values = false;
  else if (tok[3] == "complex")
    return set_errorf(GR_ERR_INVALID, "%s: complex edge values not supported", path);
  Symmetry sym;
  if (tok[4] == "general")
    sym = kGeneral;
  else if (tok[4] == "symmetric")
    sym = kSymmetric;
  else if (tok[4] == "skew-symmetric")
    sym = kSkewSymmetric;
  else if (tok[4] == "hermitian")
    sym = kHermitian;
  else
    return set_errorf(GR_ERR_INVALID, "%s: unrecognized symmetry type '%s'", path,
                      tok[4].c_str());
  if (sym == kHermitian && want_values)
    return set_errorf(GR_ERR_INVALID, "%s: hermitian edge values not supported", path);
  return load_text(f, path, 2, '%', true, true, sym, want_values, g);
}

bool read_exact(gzFile f, void* dst, size_t bytes) {
  char* p = static_cast<char*>(dst);
  while (bytes > 0) {
    const unsigned chunk = bytes > (1u << 30) ? (1u << 30) : (unsigned)bytes;
    const int got = gzread(f, p, chunk);
    if (got <= 0) return false;
    p += got;
    bytes -= (size_t)got;
  }
  return true;
}

// graphio.cpp:265-362 with expand = true: the CSR comes back as an edge list
// in CSR order.
int load_gr(gzFile f, const char* path, bool want_values, gr_graphfile* g) {
  uint64_t head[4];  // version, edge-data size, nVertices, nEdges
  if (!read_exact(f, head, sizeof head))
    return set_errorf(GR_ERR_INVALID, "%s: error reading lonestar binary CSR header", path);
  const uint64_t esz = head[1], nv = head[2], ne = head[3];
  if (esz == 0 && want_values)
    return set_errorf(GR_ERR_INVALID, "%s: file does not have edge values", path);
  if (esz != 0 && esz != 4)
    return set_errorf(GR_ERR_INVALID, "%s: file edge data is %llu bytes wide", path,
                      (unsigned long long)esz);
  if (nv >= (uint64_t)INT_MAX || ne > (uint64_t)INT_MAX)
    return set_errorf(GR_ERR_INVALID, "%s: %llu vertices / %llu edges exceed int32", path,
                      (unsigned long long)nv, (unsigned long long)ne);
  std::vector<uint64_t> ends(nv);
  g->dsts.resize(ne);
  if ((nv && !read_exact(f, ends.data(), 8 * nv)) ||
      (ne && !read_exact(f, g->dsts.data(), 4 * ne)))
    return set_errorf(GR_ERR_INVALID, "%s: truncated lonestar binary CSR file", path);
  uint64_t prev = 0;
  for (uint64_t v = 0; v < nv; ++v) {
    if (ends[v] < prev || ends[v] > ne)
      return set_errorf(GR_ERR_INVALID, "%s: offset of vertex %llu out of order", path,
                        (unsigned long long)v);
    prev = ends[v];
  }
  if (prev != ne)
    return set_errorf(GR_ERR_INVALID, "%s: offsets end at %llu, not at nEdges %llu", path,
                      (unsigned long long)prev, (unsigned long long)ne);
  for (uint64_t e = 0; e < ne; ++e)
    if ((uint32_t)g->dsts[e] >= nv)
      return set_errorf(GR_ERR_INVALID, "%s: destination of edge %llu out of range", path,
                        (unsigned long long)e);
  if (want_values) {
    if (ne % 2) {
      uint32_t pad;
      if (!read_exact(f, &pad, 4))
        return set_errorf(GR_ERR_INVALID, "%s: truncated lonestar binary CSR file", path);
    }
    g->values.resize(ne);
    if (ne && !read_exact(f, g->values.data(), 4 * ne))
      return set_errorf(GR_ERR_INVALID, "%s: truncated edge data", path);
  }
  g->srcs.resize(ne);
  uint64_t e = 0;
  for (uint64_t v = 0; v < nv; ++v)
    for (; e < ends[v]; ++e) g->srcs[e] = (int32_t)v;
  g->n_vertices = (int32_t)nv;
  g->has_values = want_values;
  return GR_OK;
}

enum Format { kEdge, kMtx, kGr };

// graphio.cpp:446-472: the extension after the last '.', skipping a final
// ".gz"; matched by prefix as there (".edges" reads as .edge, ".graph" as .gr).
int detect_format(const char* path, Format* fmt) {
  const char* dot = std::strrchr(path, '.');
  if (dot && std::strcmp(dot, ".gz") == 0) {
    const char* q = dot - 1;
    while (q >= path && *q != '.') --q;
    dot = q >= path ? q : nullptr;
  }
  if (!dot) return set_errorf(GR_ERR_INVALID, "%s: no filetype extension", path);
  if (std::strncmp(dot, ".edge", 5) == 0)
    *fmt = kEdge;
  else if (std::strncmp(dot, ".mtx", 4) == 0)
    *fmt = kMtx;
  else if (std::strncmp(dot, ".gr", 3) == 0)
    *fmt = kGr;
  else
    return set_errorf(GR_ERR_INVALID, "unrecognized filetype extension %s", dot);
  return GR_OK;
}

int load(const char* path, bool want_values, gr_graphfile* g) {
  Format fmt;
  if (int rc = detect_format(path, &fmt)) return rc;
  GzFile f;
  f.f = gzopen(path, "rb");
  if (!f.f) return set_errorf(GR_ERR_INVALID, "error opening file %s", path);
  gzbuffer(f.f, 1 << 20);
  switch (fmt) {
    case kEdge:  // graphio.cpp:172-181: '#' comments, 0-based, no values
      return load_text(f.f, path, 1, '#', false, false, kGeneral, false, g);
    case kMtx:
      return load_mtx(f.f, path, want_values, g);
    default:
      return load_gr(f.f, path, want_values, g);
  }
}

}  // namespace

extern "C" int gr_graphfile_load(const char* path, int want_values, gr_graphfile** out) {
  if (!path || !out) return set_error(GR_ERR_INVALID, "gr_graphfile_load: null argument");
  *out = nullptr;
  gr_graphfile* g = new (std::nothrow) gr_graphfile;
  if (!g) return set_error(GR_ERR_NOMEM, "gr_graphfile_load: out of host memory");
  int rc;
  try {
    rc = load(path, want_values != 0, g);
  } catch (const std::bad_alloc&) {
    rc = set_errorf(GR_ERR_NOMEM, "%s: out of host memory", path);
  }
  if (rc != GR_OK) {
    delete g;
    return rc;
  }
  *out = g;
  return GR_OK;
}

extern "C" int32_t gr_graphfile_n_vertices(const gr_graphfile* f) { return f ? f->n_vertices : 0; }
extern "C" int64_t gr_graphfile_n_edges(const gr_graphfile* f) {
  return f ? (int64_t)f->srcs.size() : 0;
}
extern "C" const int32_t* gr_graphfile_srcs(const gr_graphfile* f) {
  return f ? f->srcs.data() : nullptr;
}
extern "C" const int32_t* gr_graphfile_dsts(const gr_graphfile* f) {
  return f ? f->dsts.data() : nullptr;
}
extern "C" const int32_t* gr_graphfile_values(const gr_graphfile* f) {
  return f && f->has_values ? f->values.data() : nullptr;
}
extern "C" void gr_graphfile_free(gr_graphfile* f) { delete f; }

extern "C" int gr_write_graph_gr(const char* path, int32_t n_vertices, int32_t n_edges,
                                 const int32_t* offsets, const int32_t* dsts,
                                 const int32_t* values) {
  if (!path || n_vertices < 0 || n_edges < 0 || !offsets || (n_edges && !dsts))
    return set_error(GR_ERR_INVALID, "gr_write_graph_gr: bad argument");
  if (offsets[0] != 0 || offsets[n_vertices] != n_edges)
    return set_error(GR_ERR_INVALID, "gr_write_graph_gr: offsets must run from 0 to n_edges");
  for (int32_t v = 0; v < n_vertices; ++v)
    if (offsets[v + 1] < offsets[v])
      return set_errorf(GR_ERR_INVALID, "gr_write_graph_gr: offsets decrease at vertex %d", v);
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_errorf(GR_ERR_INVALID, "unable to write to file %s", path);
  bool ok = true;
  const uint64_t head[4] = {1, values ? 4u : 0u, (uint64_t)n_vertices, (uint64_t)n_edges};
  ok = ok && std::fwrite(head, 8, 4, f) == 4;
  std::vector<uint64_t> ends(4096);
  for (int32_t v0 = 0; ok && v0 < n_vertices; v0 += 4096) {  // end offsets, no leading 0
    const int32_t k = n_vertices - v0 < 4096 ? n_vertices - v0 : 4096;
    for (int32_t i = 0; i < k; ++i) ends[i] = (uint64_t)offsets[v0 + i + 1];
    ok = std::fwrite(ends.data(), 8, (size_t)k, f) == (size_t)k;
  }
  if (ok && n_edges) ok = std::fwrite(dsts, 4, (size_t)n_edges, f) == (size_t)n_edges;
  if (ok && n_edges % 2) {
    const uint32_t pad = 0;
    ok = std::fwrite(&pad, 4, 1, f) == 1;
  }
  if (ok && values && n_edges) ok = std::fwrite(values, 4, (size_t)n_edges, f) == (size_t)n_edges;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GR_OK : set_errorf(GR_ERR_INVALID, "error writing %s", path);
}

extern "C" int gr_write_graph_mtx(const char* path, int32_t n_vertices, int32_t n_edges,
                                  const int32_t* srcs, const int32_t* dsts,
                                  const int32_t* values) {
  if (!path || n_vertices < 0 || n_edges < 0 || (n_edges && (!srcs || !dsts)))
    return set_error(GR_ERR_INVALID, "gr_write_graph_mtx: bad argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return set_errorf(GR_ERR_INVALID, "unable to write to file %s", path);
  bool ok = std::fprintf(f, "%%%%MatrixMarket matrix coordinate Integer general\n%d %d %d\n",
                         n_vertices, n_vertices, n_edges) > 0;
  for (int32_t i = 0; ok && i < n_edges; ++i) {
    if (values)
      ok = std::fprintf(f, "%d %d %d\n", srcs[i] + 1, dsts[i] + 1, values[i]) > 0;
    else
      ok = std::fprintf(f, "%d %d\n", srcs[i] + 1, dsts[i] + 1) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GR_OK : set_errorf(GR_ERR_INVALID, "error writing %s", path);
}

extern "C" int gr_edge_list_to_csr(int32_t n_vertices, int32_t n_edges, const int32_t* srcs,
                                   const int32_t* dsts, int32_t* offsets, int32_t* out_dsts,
                                   int32_t* sort_indices) {
  if (n_vertices < 0 || n_edges < 0 || !offsets || (n_edges && !srcs) || (out_dsts && !dsts))
    return set_error(GR_ERR_INVALID, "gr_edge_list_to_csr: bad argument");
  for (int32_t i = 0; i < n_edges; ++i)
    if (srcs[i] < 0 || srcs[i] >= n_vertices)
      return set_errorf(GR_ERR_INVALID, "gr_edge_list_to_csr: source of edge %d out of range", i);
  // counting sort by source, stable in edge-list position
  std::vector<int32_t> next((size_t)n_vertices + 1, 0);
  for (int32_t i = 0; i < n_edges; ++i) ++next[(size_t)srcs[i] + 1];
  for (int32_t v = 0; v < n_vertices; ++v) next[(size_t)v + 1] += next[v];
  for (int32_t v = 0; v <= n_vertices; ++v) offsets[v] = next[v];
  for (int32_t i = 0; i < n_edges; ++i) {
    const int32_t slot = next[srcs[i]]++;
    if (out_dsts) out_dsts[slot] = dsts[i];
    if (sort_indices) sort_indices[slot] = i;
  }
  return GR_OK;
}
