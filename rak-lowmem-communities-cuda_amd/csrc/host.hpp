#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace rlc {

const char *last_error_cstr();

// offsets monotone, offsets[0] = 0, offsets[n] = m, targets < n, weights finite and >= 0.
rlc_status validate_csr(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *edg, const float *wei,
                        uint64_t *max_degree);

// Stable 3-way partition of 0..n-1 by degree: [deg < t1 | t1 <= deg < t2 | deg >= t2].
void build_class_lists(uint64_t n, const uint64_t *off, uint64_t t1, uint64_t t2, std::vector<uint32_t> &list,
                       uint64_t &c1, uint64_t &c2);

}  // namespace rlc
