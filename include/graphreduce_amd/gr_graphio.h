/*
 * graphreduce_amd graph files: the readers and writers of the reference's
 * graphio.cpp (loadGraph and friends, graphio.cpp:69-473) and the CSR
 * conversion used by its mtx2gr tool (util.cuh:79-108), behind a C ABI.
 *
 * Formats (chosen by the file extension, after an optional ".gz"):
 *  .edge  GraphLab/SNAP text: "src dst" per line, 0-based, '#' comments,
 *         self-loops dropped, no edge values (graphio.cpp:172-181).
 *  .mtx   MatrixMarket coordinate text: "%%MatrixMarket matrix coordinate
 *         <type> <symmetry>" header, '%' comments, a size line (ignored),
 *         then "row col [value]" 1-based; self-loops dropped; "symmetric"
 *         adds the reverse arc with the same value, "skew-symmetric" with the
 *         negated value, "hermitian" adds the reverse arc (values not
 *         supported there); values are read as float and truncated to int;
 *         type "pattern" has no values, "complex" is refused
 *         (graphio.cpp:184-261).
 *  .gr    Lonestar binary CSR, little endian (graphio.cpp:265-362): u64
 *         version, u64 edge-data size (4, or 0 for none), u64 nVertices, u64
 *         nEdges, u64 end offset per vertex, u32 destination per edge (padded
 *         to 8 bytes), u32 edge data per edge.  The edges come out in CSR
 *         order, self-loops kept.
 * nVertices of a text file is 1 + the largest vertex id on a data line
 * (the size line is not used), as in graphio.cpp:161-165.
 *
 * Errors return GR_ERR_INVALID (bad or unreadable file: message with the line
 * or field) or GR_ERR_NOMEM; the reference prints and exit(1)s instead.
 * Where the reference reads undefined memory (a missing value column, a
 * hermitian file with values, ids out of range, malformed .gr offsets) or
 * refuses a valid file (a .gr without edge data even when no values are
 * wanted, graphio.cpp:292-301), these functions return an error or read the
 * file, respectively.  Text headers may also be separated by tabs and end in
 * "\r\n"; a .gr file may be gzip-compressed.
 */
#ifndef GRAPHREDUCE_AMD_GR_GRAPHIO_H
#define GRAPHREDUCE_AMD_GR_GRAPHIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edge list read from a file; owned by the library until gr_graphfile_free. */
typedef struct gr_graphfile gr_graphfile;

/* loadGraph (graphio.cpp:441-473).  want_values != 0 asks for the edge
 * values (edgeValues != NULL in the reference); a text file without them
 * (.edge, MatrixMarket "pattern") then yields none and gr_graphfile_values
 * returns NULL, while a .gr file with edge-data size 0 is an error, as in
 * graphio.cpp:292-296. */
int gr_graphfile_load(const char* path, int want_values, gr_graphfile** out);
int32_t gr_graphfile_n_vertices(const gr_graphfile* f);
int64_t gr_graphfile_n_edges(const gr_graphfile* f);
const int32_t* gr_graphfile_srcs(const gr_graphfile* f);
const int32_t* gr_graphfile_dsts(const gr_graphfile* f);
const int32_t* gr_graphfile_values(const gr_graphfile* f); /* NULL: no values */
void gr_graphfile_free(gr_graphfile* f);

/* writeGraph_binaryCSR (graphio.cpp:365-413): offsets[n_vertices + 1] and
 * dsts[n_edges] of a CSR; values may be NULL (edge-data size 0). */
int gr_write_graph_gr(const char* path, int32_t n_vertices, int32_t n_edges,
                      const int32_t* offsets, const int32_t* dsts, const int32_t* values);

/* writeGraph_mtx (graphio.cpp:416-438): "%%MatrixMarket matrix coordinate
 * Integer general", the size line, then 1-based "src dst [value]" lines in
 * the given order; values may be NULL. */
int gr_write_graph_mtx(const char* path, int32_t n_vertices, int32_t n_edges,
                       const int32_t* srcs, const int32_t* dsts, const int32_t* values);

/* edgeListToCSR (util.cuh:79-108): offsets[n_vertices + 1], out_dsts[n_edges]
 * (may be NULL) and sort_indices[n_edges] (may be NULL; edge-list position of
 * each CSR slot).  Edges are ordered by source, and by edge-list position
 * within a source (the reference's std::sort leaves that order unspecified). */
int gr_edge_list_to_csr(int32_t n_vertices, int32_t n_edges, const int32_t* srcs,
                        const int32_t* dsts, int32_t* offsets, int32_t* out_dsts,
                        int32_t* sort_indices);

#ifdef __cplusplus
}
#endif

#endif
