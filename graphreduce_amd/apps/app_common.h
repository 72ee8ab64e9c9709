// Shared pieces of the app drivers: command line, graph input, timing.
//
// Command line as in the reference apps (util.cu:33-100, format "si-t-d-m|s"
// for bfs/sssp, "s-t-d|s" for pagerank/cc): positional arguments in order,
// single-letter flags anywhere.  Graph input (graphio.cpp:441-473 in the
// reference reads .mtx/.edge/.gr files) accepts:
//   rmat:SCALE:SEED[:sym]   RMAT, edge factor 16, a/b/c = .57/.19/.19,
//                           self-loops and duplicates removed
//   grid:SIDE:SEED          4-neighbour grid, 1 % of edges deleted, both arcs
//   FILE.edge / FILE.mtx / FILE.gr [.gz]
//                           graph files as graphio.cpp:441-473 reads them
//                           (gr_graphio.h)
// Synthetic inputs get hash weights uniform on [1, 64] (seed = SEED).
#ifndef GRAPHREDUCE_AMD_APP_COMMON_H
#define GRAPHREDUCE_AMD_APP_COMMON_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <string>
#include <vector>

#include "graphreduce_amd/gr_abi.h"
#include "graphreduce_amd/gr_graphio.h"

namespace app {

inline int64_t currentTime() {  // microseconds, util.cu:25-30
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return (int64_t)tv.tv_sec * 1000000 + tv.tv_usec;
}

struct Args {
  std::vector<std::string> pos;
  bool t = false, d = false, m = false;
};

inline Args parse(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    if (argv[i][0] == '-' && argv[i][1] && !argv[i][2]) {
      switch (argv[i][1]) {
        case 't': a.t = true; break;
        case 'd': a.d = true; break;
        case 'm': a.m = true; break;
        default: fprintf(stderr, "unknown flag %s\n", argv[i]); exit(1);
      }
    } else {
      a.pos.push_back(argv[i]);
    }
  }
  return a;
}

inline void die(int rc, const char* what) {
  if (rc != GR_OK) {
    fprintf(stderr, "%s: %s\n", what, gr_last_error());
    exit(1);
  }
}

inline std::vector<std::string> split(const std::string& s, char c) {
  std::vector<std::string> out;
  size_t b = 0;
  for (;;) {
    size_t e = s.find(c, b);
    out.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
    if (e == std::string::npos) break;
    b = e + 1;
  }
  return out;
}

// Reads a graph; edgeValues (may be null) receives one int weight per arc.
inline void loadGraph(const std::string& name, int& nVertices, std::vector<int>& srcs,
                      std::vector<int>& dsts, std::vector<int>* edgeValues) {
  srcs.clear();
  dsts.clear();
  if (edgeValues) edgeValues->clear();
  std::vector<std::string> f = split(name, ':');
  if (f[0] == "rmat" && f.size() >= 3) {
    const int scale = atoi(f[1].c_str());
    const uint64_t seed = strtoull(f[2].c_str(), nullptr, 10);
    const int sym = f.size() > 3 && f[3] == "sym";
    int64_t cap = ((int64_t)16 << scale) * (sym ? 2 : 1), n = 0;
    srcs.resize(cap);
    dsts.resize(cap);
    die(gr_host_rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed, sym, srcs.data(), dsts.data(), cap,
                           &n),
        "rmat");
    srcs.resize(n);
    dsts.resize(n);
    nVertices = 1 << scale;
    if (edgeValues) {
      edgeValues->resize(n);
      die(gr_host_hash_weights(seed, n, 1, 6, edgeValues->data()), "weights");
    }
    return;
  }
  if (f[0] == "grid" && f.size() >= 3) {
    const int side = atoi(f[1].c_str());
    const uint64_t seed = strtoull(f[2].c_str(), nullptr, 10);
    int64_t n = 0;
    die(gr_host_grid_edges(side, 0.01, seed, nullptr, nullptr, 0, &n), "grid");
    srcs.resize(n);
    dsts.resize(n);
    die(gr_host_grid_edges(side, 0.01, seed, srcs.data(), dsts.data(), n, &n), "grid");
    nVertices = side * side;
    if (edgeValues) {
      edgeValues->resize(n);
      die(gr_host_hash_weights(seed, n, 1, 6, edgeValues->data()), "weights");
    }
    return;
  }
  // graph file: .edge / .mtx / .gr, optionally .gz (gr_graphio.h)
  gr_graphfile* gf = nullptr;
  die(gr_graphfile_load(name.c_str(), edgeValues ? 1 : 0, &gf), "loadGraph");
  const int64_t m = gr_graphfile_n_edges(gf);
  nVertices = gr_graphfile_n_vertices(gf);
  srcs.assign(gr_graphfile_srcs(gf), gr_graphfile_srcs(gf) + m);
  dsts.assign(gr_graphfile_dsts(gf), gr_graphfile_dsts(gf) + m);
  if (edgeValues && gr_graphfile_values(gf))
    edgeValues->assign(gr_graphfile_values(gf), gr_graphfile_values(gf) + m);
  gr_graphfile_free(gf);
}

// Vertex of maximum out-degree, first on ties (bfs.cu:175-195).
inline int maxOutDegreeVertex(int nVertices, const std::vector<int>& srcs, int* degree) {
  std::vector<int> deg(nVertices, 0);
  for (int s : srcs) deg[s]++;
  int best = -1, bestDeg = -1;
  for (int i = 0; i < nVertices; ++i)
    if (deg[i] > bestDeg) {
      bestDeg = deg[i];
      best = i;
    }
  if (degree) *degree = bestDeg;
  return best;
}

inline std::vector<int> outDegrees(int nVertices, const std::vector<int>& srcs) {
  std::vector<int> deg(nVertices, 0);
  for (int s : srcs) deg[s]++;
  return deg;
}

}  // namespace app

#endif
