// gr2mtx: any graph file loadGraph reads -> MatrixMarket text, as the
// reference's gr2mtx.cpp:22-46 (edges in file order, values if the file has
// them; a .gr without edge data, which the reference cannot read, is written
// without a value column).
#include <stdio.h>
#include <stdlib.h>

#include "graphreduce_amd/gr_abi.h"
#include "graphreduce_amd/gr_graphio.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    printf("Usage: gr2mtx input output\n");
    return 1;
  }
  gr_graphfile* f = nullptr;
  if (gr_graphfile_load(argv[1], 1, &f) != GR_OK &&
      gr_graphfile_load(argv[1], 0, &f) != GR_OK) {
    fprintf(stderr, "%s\n", gr_last_error());
    return 1;
  }
  const int32_t nV = gr_graphfile_n_vertices(f);
  const int64_t nE = gr_graphfile_n_edges(f);
  printf("Read input file with %d vertices and %lld edges\n", nV, (long long)nE);
  printf("writing output\n");
  if (gr_write_graph_mtx(argv[2], nV, (int32_t)nE, gr_graphfile_srcs(f), gr_graphfile_dsts(f),
                         gr_graphfile_values(f)) != GR_OK) {
    fprintf(stderr, "%s\n", gr_last_error());
    return 1;
  }
  gr_graphfile_free(f);
  return 0;
}
