// Counter-based random streams shared by the host and device graph generators.
//
// Every random draw is a pure function of (seed, stream, index), so the host
// generator (used by the tests and the CPU checker) and the device generator
// (used for the large benchmark graphs) produce bit-identical edge lists no
// matter how the work is split over threads or workgroups.
//
// The mixing function is the splitmix64 finaliser (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014).
#ifndef GRAPHREDUCE_AMD_GR_HASH_H
#define GRAPHREDUCE_AMD_GR_HASH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GR_HD __host__ __device__ __forceinline__
#else
#define GR_HD static inline
#endif

namespace gr {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

// Domain-separation tags: one independent stream per use of a seed.
constexpr uint64_t kStreamRmat   = 0x524D4154ULL;  // "RMAT"
constexpr uint64_t kStreamWeight = 0x57474854ULL;  // "WGHT"
constexpr uint64_t kStreamGrid   = 0x47524944ULL;  // "GRID"
constexpr uint64_t kStreamSource = 0x53524345ULL;  // "SRCE"

GR_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Base of the stream (seed, tag).
GR_HD uint64_t stream_base(uint64_t seed, uint64_t tag) {
  return mix64(mix64(seed ^ (tag * kGolden)) + tag);
}

// Draw number `index` of a stream.
GR_HD uint64_t stream_draw(uint64_t base, uint64_t index) {
  return mix64(base + (index + 1) * kGolden);
}

// RMAT quadrant thresholds as 32-bit fixed-point fractions: a draw u (uniform
// on [0, 2^32)) picks quadrant a if u < ta, b if u < tb, c if u < tc, else d.
struct RmatThresholds {
  uint32_t ta, tb, tc;
};

static inline RmatThresholds rmat_thresholds(double a, double b, double c) {
  const double s = 4294967296.0;
  RmatThresholds t;
  double ca = a * s, cb = (a + b) * s, cc = (a + b + c) * s;
  t.ta = (uint32_t)(ca >= s ? 4294967295.0 : ca);
  t.tb = (uint32_t)(cb >= s ? 4294967295.0 : cb);
  t.tc = (uint32_t)(cc >= s ? 4294967295.0 : cc);
  return t;
}

// Edge `k` of an RMAT graph with 2^scale vertices.  Level l (0 = most
// significant bit) uses 32-bit half (l & 1) of draw (k * words + l / 2).
// Quadrants: a = (0,0), b = (0,1) sets the dst bit, c = (1,0) sets the src
// bit, d = (1,1).
GR_HD void rmat_edge(uint64_t base, uint64_t k, int scale, RmatThresholds t,
                     uint32_t* src, uint32_t* dst) {
  const int words = (scale + 1) / 2;
  uint32_t s = 0, d = 0;
  uint64_t w = 0;
  for (int l = 0; l < scale; ++l) {
    if ((l & 1) == 0) w = stream_draw(base, k * (uint64_t)words + (uint64_t)(l >> 1));
    uint32_t u = (l & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
    uint32_t bit = 1u << (scale - 1 - l);
    if (u < t.ta) {
    } else if (u < t.tb) {
      d |= bit;
    } else if (u < t.tc) {
      s |= bit;
    } else {
      s |= bit;
      d |= bit;
    }
  }
  *src = s;
  *dst = d;
}

// Edge weight of edge-list position i: uniform on [lo, lo + span) with span a
// power of two (top bits of the draw, so exactly uniform).
GR_HD uint32_t hash_weight(uint64_t base, uint64_t i, uint32_t lo, int span_log2) {
  return lo + (uint32_t)(stream_draw(base, i) >> (64 - span_log2));
}

// Grid edge deletion: undirected edge e is kept when its draw is at or above
// the threshold floor(frac * 2^64).
GR_HD bool grid_keep(uint64_t base, uint64_t e, uint64_t threshold) {
  return stream_draw(base, e) >= threshold;
}

}  // namespace gr

#endif  // GRAPHREDUCE_AMD_GR_HASH_H
