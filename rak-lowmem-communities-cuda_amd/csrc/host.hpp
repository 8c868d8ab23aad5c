#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace rlc {

const char *last_error_cstr();

// offsets monotone, offsets[0] = 0, offsets[n] = m, targets < n, weights finite and >= 0.
rlc_status validate_csr(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *edg, const float *wei,
                        uint64_t *max_degree);

// Vertex order of one iteration: grouped by degree class (0: deg < t1, 1: t1 <= deg < t2,
// 2: deg >= t2), then by sub-sweep (u / 64) % sweeps, ascending id inside each group.
// seg[c * sweeps + s] is the first position of (class c, sub-sweep s); seg[3 * sweeps] = n.
void build_class_lists(uint64_t n, const uint64_t *off, uint64_t t1, uint64_t t2, int sweeps,
                       std::vector<uint32_t> &list, std::vector<uint64_t> &seg);

}  // namespace rlc
