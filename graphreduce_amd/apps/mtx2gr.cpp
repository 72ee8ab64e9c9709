// mtx2gr: any graph file loadGraph reads -> Lonestar binary CSR (.gr), as
// the reference's mtx2gr.cpp:22-53: read with edge values, CSR by source,
// values permuted with the edges.  A file without values (MatrixMarket
// "pattern", .edge) gives a .gr with edge-data size 0 (the reference indexes
// an empty vector there).
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "graphreduce_amd/gr_abi.h"
#include "graphreduce_amd/gr_graphio.h"

static void check(int rc) {
  if (rc != GR_OK) {
    fprintf(stderr, "%s\n", gr_last_error());
    exit(1);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    printf("Usage: mtx2gr input output\n");
    return 1;
  }
  gr_graphfile* f = nullptr;
  check(gr_graphfile_load(argv[1], 1, &f));
  const int32_t nV = gr_graphfile_n_vertices(f);
  const int64_t nE = gr_graphfile_n_edges(f);
  printf("Read input file with %d vertices and %lld edges\n", nV, (long long)nE);
  printf("Converting to CSR\n");
  std::vector<int32_t> offsets((size_t)nV + 1), dsts((size_t)nE), idx((size_t)nE);
  check(gr_edge_list_to_csr(nV, (int32_t)nE, gr_graphfile_srcs(f), gr_graphfile_dsts(f),
                            offsets.data(), dsts.data(), idx.data()));
  const int32_t* vals = gr_graphfile_values(f);
  std::vector<int32_t> sorted;
  if (vals) {
    sorted.resize((size_t)nE);
    for (int64_t i = 0; i < nE; ++i) sorted[i] = vals[idx[i]];
  }
  printf("writing output\n");
  check(gr_write_graph_gr(argv[2], nV, (int32_t)nE, offsets.data(), dsts.data(),
                          vals ? sorted.data() : nullptr));
  gr_graphfile_free(f);
  return 0;
}
