#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace rlc {

const char *last_error_cstr();

// offsets monotone, offsets[0] = 0, offsets[n] = m, targets < n, weights finite and >= 0.
rlc_status validate_csr(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *edg, const float *wei,
                        uint64_t *max_degree);

// Sub-sweep segment of vertex u: its 64-vertex block id, round-robin over S segments, so that
// a run of consecutive ids (often one community) is spread over all segments and later
// segments see the labels the earlier ones wrote.
uint64_t segment_of(uint64_t u, uint64_t S);

// Device layout of a graph: vertices split into `segments` sub-sweep segments by
// segment_of(id), stored segment after segment, by degree descending inside a segment
// (ties: ascending id), so that the most referenced labels share cache lines and every degree
// class is a contiguous range of each segment. perm[p] = original id of the vertex at position
// p, inv[u] = position of original vertex u, seg[s] = first position of segment s (seg[S] = n).
// Each row keeps its original edge order; targets are renumbered to positions.
struct Layout {
  std::vector<uint32_t> perm, inv;
  std::vector<uint64_t> seg;
  std::vector<uint64_t> off;
  std::vector<uint32_t> edg;
  std::vector<float> wei;  // empty for unit weights
};
rlc_status build_layout(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *edg, const float *wei,
                        int segments, Layout &out);

}  // namespace rlc
