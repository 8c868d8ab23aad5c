// Internal interface of the device-side graph construction: CSR from an edge list (sorted,
// deduplicated, optionally symmetrised) and the degree-descending segmented layout.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "rlc.h"

namespace rlc {

// A CSR on the device (arrays owned by the caller after a successful call; wei may be null).
struct DevCsr {
  uint64_t n = 0, m = 0;
  uint64_t *off = nullptr;
  uint32_t *edg = nullptr;
  float *wei = nullptr;
  void release();
};

// The library's layout of a graph: vertices split into sub-sweep segments by segment_of(id),
// stored segment after segment, by degree descending inside a segment (ties: ascending id), so
// that the most referenced labels share cache lines and every degree class is a contiguous range
// of each segment. Each row keeps its edge order; targets are renumbered to positions. Device
// arrays plus the host copies the run loop needs: h_off, h_perm[p] = caller id at position p,
// h_seg[s] = first position of segment s (h_seg[S] = n).
struct DevLayout {
  DevCsr csr;
  uint32_t *perm = nullptr;  // position -> caller id (device)
  uint64_t max_degree = 0;
  std::vector<uint64_t> h_off;
  std::vector<uint32_t> h_perm;
  std::vector<uint64_t> h_seg;
};

// Layout of the CSR `in` (device arrays, not modified) with `segments` sub-sweep segments.
// Allocates out.csr and out.perm.
rlc_status device_layout(hipStream_t st, const DevCsr &in, int segments, DevLayout &out);

// CSR of n vertices from `count` device edges (src[i], dst[i], w[i]; w may be null: weight 1).
// Adjacency lists ascending; a repeated (u, v) keeps the weight of its last occurrence. With
// symmetrize, every (u, v) also yields (v, u) unless (v, u) is present itself (then that one's
// weight is kept). Allocates out.
rlc_status device_csr_from_edges(hipStream_t st, uint64_t n, uint64_t count, const uint32_t *src, const uint32_t *dst,
                                 const float *w, bool symmetrize, DevCsr &out);

}  // namespace rlc
