// Internal interface between the C ABI (api.hip) and the kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rlc.h"

namespace rlc {

// CSR in HBM. wei == nullptr: unit weights (no weight stream is read).
struct DevGraph {
  const uint64_t *off;
  const uint32_t *edg;
  const float *wei;
  uint64_t n;
  uint64_t m;
};

// Iteration bookkeeping kept on the device, so that the loop runs without host round trips.
struct DevCounters {
  unsigned long long A, E, C, X;  // current iteration: active, their edges, changed, their edges
  unsigned long long nmoves;      // synchronous mode: length of the change list
  unsigned int iter;              // iterations completed
  unsigned int done;              // set once converged or at the iteration cap
  unsigned long long last_changed;
  unsigned long long hist[RLC_MAX_ITER_STATS][4];
};

struct DevState {
  uint32_t *lab;     // labels (community of each vertex)
  uint8_t *aff;      // vertex-pruning flags: 1 = process in the next iteration
  uint64_t *moves;   // synchronous mode change list: (u << 32) | new label
  DevCounters *ctr;
};

// Vertex order of an iteration: list[seg[c * sweeps + s] .. seg[c * sweeps + s + 1]) holds the
// vertices of degree class c (0 lane-per-vertex, 1 wave-per-vertex, 2 workgroup-per-vertex) in
// sub-sweep s. Sub-sweeps run one after the other; inside one, the three classes run in turn.
struct ClassLists {
  const uint32_t *list;
  const uint64_t *seg;  // host array of 3 * sweeps + 1 entries
  int sweeps;
};

struct LoopParams {
  bool sync;
  bool mg;          // MG8 (true) or BM (false)
  int pickstep;
  double tolerance;
  int max_iterations;
  int block_parts;  // parts of the workgroup class: 256
};

// labels = init ? init : ids; flags = 1; counters reset (done = max_iterations <= 0)
void launch_init_labels(const DevGraph &g, const DevState &s, const uint32_t *init, hipStream_t st);
void launch_mark_all(const DevGraph &g, const DevState &s, int max_iterations, hipStream_t st);
// One iteration: move launches for each non-empty class, the apply pass (synchronous mode) and
// the finish step. ev_move0/ev_move1 (may be null) bracket the move launches.
void launch_iteration(const DevGraph &g, const DevState &s, const ClassLists &cl, const LoopParams &p,
                      hipStream_t st, hipEvent_t ev_move0, hipEvent_t ev_move1);
// FP64 modularity terms: out[0] = sum of intra-community weight, out[1] = total weight,
// out[2] = sum over labels of tot^2, out[3] = number of distinct labels. Needs labels < n.
// scratch: n doubles + n bytes.
void launch_modularity(const DevGraph &g, const uint32_t *lab, double *ctot, uint8_t *seen, double *out,
                       hipStream_t st);

}  // namespace rlc
