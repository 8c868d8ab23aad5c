// Host-side helpers (g++ / OpenMP): error state, CSR validation, degree-class lists.
#include "host.hpp"

#include <omp.h>

#include <algorithm>
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

rlc_status build_layout(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *edg, const float *wei,
                        int segments, Layout &L) {
  const uint64_t S = (uint64_t)(segments < 1 ? 1 : segments);
  try {
    L.perm.resize(n);
    L.inv.resize(n);
    L.off.resize(n + 1);
    L.edg.resize(m);
    if (wei) L.wei.resize(m);
  } catch (...) {
    return fail(RLC_ERR_OUT_OF_MEMORY, "graph layout: host allocation failed");
  }
  uint64_t maxd = 0;
  for (uint64_t u = 0; u < n; ++u) maxd = std::max<uint64_t>(maxd, off[u + 1] - off[u]);
  // counting sort by (segment, degree descending), stable in id
  const uint64_t D = maxd + 1;
  auto key = [&](uint64_t u) { return segment_of(u, S) * D + (maxd - (off[u + 1] - off[u])); };
  std::vector<uint64_t> cnt(S * D + 1, 0);
  for (uint64_t u = 0; u < n; ++u) ++cnt[key(u) + 1];
  for (uint64_t k = 0; k < S * D; ++k) cnt[k + 1] += cnt[k];
  L.seg.resize(S + 1);
  for (uint64_t s = 0; s <= S; ++s) L.seg[s] = cnt[s * D];
  for (uint64_t u = 0; u < n; ++u) L.perm[cnt[key(u)]++] = (uint32_t)u;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < (int64_t)n; ++p) L.inv[L.perm[p]] = (uint32_t)p;
  L.off[0] = 0;
  for (uint64_t p = 0; p < n; ++p) {
    const uint64_t u = L.perm[p];
    L.off[p + 1] = L.off[p] + (off[u + 1] - off[u]);
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p = 0; p < (int64_t)n; ++p) {
    const uint64_t u = L.perm[p], b = off[u], d = off[u + 1] - b, o = L.off[p];
    for (uint64_t j = 0; j < d; ++j) L.edg[o + j] = L.inv[edg[b + j]];
    if (wei)
      for (uint64_t j = 0; j < d; ++j) L.wei[o + j] = wei[b + j];
  }
  return RLC_OK;
}

}  // namespace rlc
