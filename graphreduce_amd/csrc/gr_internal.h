// Library-internal declarations shared by the host and device sources.
#ifndef GRAPHREDUCE_AMD_GR_INTERNAL_H
#define GRAPHREDUCE_AMD_GR_INTERNAL_H

#include <stdint.h>

#include <vector>

#include "../../include/graphreduce_amd/gr_abi.h"

namespace gr {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const char* msg);
int set_errorf(int code, const char* fmt, ...);

// Sorted, de-duplicated RMAT keys src << scale | dst (host generator).
int64_t host_rmat_keys(int scale, int edge_factor, double a, double b, double c,
                       uint64_t seed, int symmetrize, std::vector<uint64_t>& keys);

// Grid arcs; with srcs == nullptr only counts.
int64_t host_grid_arcs(int side, double delete_frac, uint64_t seed, int32_t* srcs,
                       int32_t* dsts);

}  // namespace gr

#endif
