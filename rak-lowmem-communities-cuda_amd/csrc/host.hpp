#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace rlc {

const char *last_error_cstr();

// offsets monotone, offsets[0] = 0, offsets[n] = m, targets < n, weights finite and >= 0.
rlc_status validate_csr(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *edg, const float *wei,
                        uint64_t *max_degree);

// Sub-sweep segment of vertex u: its window of kSegmentWindow consecutive ids, round-robin over
// S segments, so that a run of consecutive ids (often one community, or a row of a lattice) is
// spread over all segments and later segments see the labels the earlier ones wrote. Windows of
// 8 measured better than 64 or 1 on a road-like grid (modularity 0.686 vs 0.643 / 0.655, in 10
// iterations instead of 18 / 20) and on SBM graphs, at equal speed.
constexpr uint64_t kSegmentWindow = 8;
uint64_t segment_of(uint64_t u, uint64_t S);

// memcpy with up to `threads` OpenMP threads for large copies (pieces of at least 1 MiB).
void parallel_memcpy(void *dst, const void *src, uint64_t bytes, int threads);

}  // namespace rlc
