// Host-side helpers (g++ / OpenMP): error state, CSR validation, degree-class lists.
#include "host.hpp"

#include <omp.h>

#include <algorithm>
#include <cstring>
#include <string>

namespace rlc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
const char *last_error_cstr() { return g_last_error.c_str(); }

rlc_status validate_csr(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *edg, const float *wei,
                        uint64_t *max_degree) {
  if (n >= 0xFFFFFFFFull) return fail(RLC_ERR_INVALID_GRAPH, "graph: n must be below 2^32-1 (label 0xFFFFFFFF is reserved)");
  if (!off) return fail(RLC_ERR_INVALID_ARGUMENT, "graph: offsets is null");
  if (m > 0 && !edg) return fail(RLC_ERR_INVALID_ARGUMENT, "graph: edges is null");
  if (off[0] != 0) return fail(RLC_ERR_INVALID_GRAPH, "graph: offsets[0] != 0");
  if (off[n] != m) return fail(RLC_ERR_INVALID_GRAPH, "graph: offsets[n] != m");
  int bad_off = 0, bad_edge = 0, bad_w = 0;
  uint64_t md = 0;
#pragma omp parallel for schedule(static) reduction(| : bad_off) reduction(max : md)
  for (int64_t u = 0; u < (int64_t)n; ++u) {
    if (off[u + 1] < off[u]) { bad_off = 1; continue; }
    uint64_t d = off[u + 1] - off[u];
    md = d > md ? d : md;
  }
  if (bad_off) return fail(RLC_ERR_INVALID_GRAPH, "graph: offsets are not non-decreasing");
#pragma omp parallel for schedule(static) reduction(| : bad_edge, bad_w)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    bad_edge |= edg[i] >= n;
    if (wei) bad_w |= !(wei[i] >= 0.0f) || !(wei[i] < __builtin_huge_valf());
  }
  if (bad_edge) return fail(RLC_ERR_INVALID_GRAPH, "graph: an edge target is >= n");
  if (bad_w) return fail(RLC_ERR_INVALID_GRAPH, "graph: weights must be finite and non-negative");
  if (max_degree) *max_degree = md;
  return RLC_OK;
}

uint64_t segment_of(uint64_t u, uint64_t S) { return S <= 1 ? 0 : (u / kSegmentWindow) % S; }

void parallel_memcpy(void *dst, const void *src, uint64_t bytes, int threads) {
  constexpr uint64_t kPiece = 1ull << 20;
  const uint64_t pieces = (bytes + kPiece - 1) / kPiece;
  if (threads <= 1 || pieces < 4) {
    std::memcpy(dst, src, bytes);
    return;
  }
  char *d = (char *)dst;
  const char *s = (const char *)src;
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int64_t k = 0; k < (int64_t)pieces; ++k) {
    const uint64_t a = (uint64_t)k * kPiece, len = std::min(kPiece, bytes - a);
    std::memcpy(d + a, s + a, len);
  }
}

}  // namespace rlc
