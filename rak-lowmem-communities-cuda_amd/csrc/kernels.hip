// HIP kernels for gfx950 (wave64): one label-propagation move iteration with Misra-Gries (8
// slots) or Boyer-Moore sketches, in three granularities chosen by vertex degree, plus the
// apply / finish / initialisation steps and the FP64 modularity reduction.
//
// Reference counterparts (inc/rakLowmemCuda.hxx): the move kernels replace
// rakLowmemMgMoveIterationGroupCukU (:269-317), rakLowmemMgMoveIterationBlockCukU (:362-419),
// rakLowmemBmMoveIterationThreadCukU (:138-161) and rakLowmemBmMoveIterationBlockCukU (:198-229);
// the finish step replaces the host loop test of rakLowmemLoopCuU (:469-490) on the device.
//
// Semantics shared by every kernel and by both modes, per processed vertex u with label d:
//   scan    feed (label[v], w) of every edge (v, w), v != u, in CSR order into the sketch
//           (lane kernel: one sequential sketch; wave / workgroup kernels: P strided parts,
//           part p scans positions p, p+P, .., merged by the fixed binary tree)
//   rescan  (MG8) exact weights of the 8 surviving labels (tree-ordered sums for P parts)
//   choose  first slot of largest weight; accept c if w > 0, c != d and, on "pick-less"
//           iterations (l % pickstep == 0), c < d (rakLowmemCuda.hxx:306-309)
//   flags   aff[u] = 0; if u moved, aff[v] = 1 for its neighbours v != u
// Asynchronous mode writes label[u] in place (as the reference does); synchronous mode reads a
// snapshot (labels are not written during the move launches) and appends (u, c) to a change
// list that the apply pass commits after the iteration.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lpa.hpp"
#include "sketch.hpp"

namespace rlc {

namespace {

constexpr int kBlock = 256;

// Per-block reduction of the four iteration counters into the device totals.
struct BlockStats {
  unsigned long long a, e, c, x;
};

__device__ __forceinline__ void flush_stats(DevCounters *ctr, unsigned long long a, unsigned long long e,
                                            unsigned long long c, unsigned long long x) {
  __shared__ unsigned long long red[4][kBlock / 64];
  a = wave_sum_u64(a);
  e = wave_sum_u64(e);
  c = wave_sum_u64(c);
  x = wave_sum_u64(x);
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) { red[0][wave] = a; red[1][wave] = e; red[2][wave] = c; red[3][wave] = x; }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[threadIdx.x][w];
    if (t) atomicAdd(&ctr->A + threadIdx.x, t);
  }
}

__device__ __forceinline__ float edge_w(const DevGraph &g, uint64_t i) { return g.wei ? g.wei[i] : 1.0f; }

// ---- lane per vertex: one sequential sketch per lane ----------------------------------------
template <bool SYNC, bool MG>
__global__ __launch_bounds__(kBlock) void k_move_lane(DevGraph g, DevState s, const uint32_t *__restrict__ list,
                                                      uint64_t count, int pickstep) {
  if (s.ctr->done) return;
  const uint32_t iter = s.ctr->iter;
  const bool pickless = pickstep > 0 && (iter % (uint32_t)pickstep) == 0;
  unsigned long long sa = 0, se = 0, sc = 0, sx = 0;
  __shared__ unsigned int wcount[kBlock / 64];
  __shared__ unsigned long long bbase;
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < count; base += (uint64_t)gridDim.x * kBlock) {
    const uint64_t idx = base + threadIdx.x;
    const bool valid = idx < count;
    const uint32_t u = valid ? list[idx] : 0u;
    const bool act = valid && s.aff[u] != 0;
    uint64_t deg = 0;
    bool moved = false;
    uint32_t cb = 0;
    if (act) {
      const uint64_t b = g.off[u], e = g.off[(uint64_t)u + 1];
      deg = e - b;
      const uint32_t d = s.lab[u];
      float wb = 0.0f;
      if (MG) {
        MG8 sk;
        sk.clear();
        for (uint64_t i = b; i < e; ++i) {
          const uint32_t v = g.edg[i];
          if (v == u) continue;
          sk.accumulate(s.lab[v], edge_w(g, i));
        }
        float acc[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) acc[j] = 0.0f;
        for (uint64_t i = b; i < e; ++i) {
          const uint32_t v = g.edg[i];
          if (v == u) continue;
          const uint32_t c = s.lab[v];
          const float w = edge_w(g, i);
#pragma unroll
          for (int j = 0; j < kSlots; ++j) acc[j] = sk.c[j] == c ? acc[j] + w : acc[j];
        }
        argmax8(sk.c, acc, cb, wb);
      } else {
        uint32_t mc = kEmpty;
        float mw = 0.0f;
        for (uint64_t i = b; i < e; ++i) {
          const uint32_t v = g.edg[i];
          if (v == u) continue;
          bm_step(mc, mw, s.lab[v], edge_w(g, i));
        }
        cb = mc;
        wb = mw;
      }
      moved = wb > 0.0f && cb != d && !(pickless && cb > d);
      s.aff[u] = 0;
      if (!SYNC && moved) {
        s.lab[u] = cb;
        for (uint64_t i = b; i < e; ++i) {
          const uint32_t v = g.edg[i];
          if (v != u) s.aff[v] = 1;
        }
      }
    }
    sa += act ? 1 : 0;
    se += deg;
    if (SYNC) {
      // block-aggregated append to the change list: one atomic per block round
      const unsigned long long mask = __ballot(moved);
      if (lane == 0) wcount[wave] = (unsigned int)__popcll(mask);
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += wcount[w];
        bbase = t ? atomicAdd(&s.ctr->nmoves, (unsigned long long)t) : 0ull;
      }
      __syncthreads();
      if (moved) {
        unsigned long long pos = bbase;
        for (int w = 0; w < wave; ++w) pos += wcount[w];
        pos += lanes_below(mask);
        s.moves[pos] = ((uint64_t)u << 32) | cb;
      }
      __syncthreads();
    } else {
      sc += moved ? 1 : 0;
      sx += moved ? deg : 0;
    }
  }
  flush_stats(s.ctr, sa, se, sc, sx);
}

// ---- P lanes per vertex (P = 64: one wave; P = 256: four waves of one workgroup) -----------
template <bool SYNC, bool MG, int P>
__global__ __launch_bounds__(kBlock) void k_move_group(DevGraph g, DevState s, const uint32_t *__restrict__ list,
                                                       uint64_t count, int pickstep) {
  constexpr int GW = P / 64;     // waves per vertex
  constexpr int VPB = kBlock / P;  // vertices per block round
  if (s.ctr->done) return;
  const uint32_t iter = s.ctr->iter;
  const bool pickless = pickstep > 0 && (iter % (uint32_t)pickstep) == 0;
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int grp = wave / GW;  // vertex slot in the block
  const int gw = wave % GW;   // wave inside the group
  const int gl = (int)threadIdx.x - grp * P;  // part index 0..P-1
  __shared__ uint32_t sh_c[GW][kSlots];
  __shared__ float sh_w[GW][kSlots];
  __shared__ uint32_t sh_cand[kSlots];
  __shared__ float sh_tot[kSlots];
  __shared__ uint32_t sh_flag;
  unsigned long long sa = 0, se = 0, sc = 0, sx = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * VPB; base < count; base += (uint64_t)gridDim.x * VPB) {
    const uint64_t idx = base + grp;
    const bool valid = idx < count;
    const uint32_t u = valid ? list[idx] : 0u;
    bool act;
    if (GW == 1) {
      act = valid && uniform_u32(s.aff[u]) != 0;
    } else {
      if (threadIdx.x == 0) sh_flag = valid ? s.aff[u] : 0u;
      __syncthreads();
      act = sh_flag != 0;
    }
    if (act) {
      const uint64_t b = g.off[u], e = g.off[(uint64_t)u + 1];
      const uint64_t deg = e - b;
      const uint32_t d = s.lab[u];
      uint32_t cb;
      float wb;
      if (MG) {
        MG8 sk;
        sk.clear();
        for (uint64_t i = b + gl; i < e; i += P) {
          const uint32_t v = g.edg[i];
          if (v == u) continue;
          sk.accumulate(s.lab[v], edge_w(g, i));
        }
        wave_merge_tree(sk);
        uint32_t cand[kSlots];
        if (GW == 1) {
#pragma unroll
          for (int j = 0; j < kSlots; ++j) cand[j] = __shfl(sk.c[j], 0);
        } else {
          if (lane == 0) {
#pragma unroll
            for (int j = 0; j < kSlots; ++j) { sh_c[gw][j] = sk.c[j]; sh_w[gw][j] = sk.w[j]; }
          }
          __syncthreads();
          if (threadIdx.x == 0) {
            // continue the binary tree across the waves: wave q absorbs wave q+k
            for (int k = 1; k < GW; k <<= 1)
              for (int q = 0; q + k < GW; q += 2 * k) {
                MG8 a;
                uint32_t oc[kSlots];
                float ow[kSlots];
                for (int j = 0; j < kSlots; ++j) {
                  a.c[j] = sh_c[q][j]; a.w[j] = sh_w[q][j];
                  oc[j] = sh_c[q + k][j]; ow[j] = sh_w[q + k][j];
                }
                a.absorb(oc, ow);
                for (int j = 0; j < kSlots; ++j) { sh_c[q][j] = a.c[j]; sh_w[q][j] = a.w[j]; }
              }
            for (int j = 0; j < kSlots; ++j) sh_cand[j] = sh_c[0][j];
          }
          __syncthreads();
#pragma unroll
          for (int j = 0; j < kSlots; ++j) cand[j] = sh_cand[j];
        }
        float acc[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) acc[j] = 0.0f;
        for (uint64_t i = b + gl; i < e; i += P) {
          const uint32_t v = g.edg[i];
          if (v == u) continue;
          const uint32_t c = s.lab[v];
          const float w = edge_w(g, i);
#pragma unroll
          for (int j = 0; j < kSlots; ++j) acc[j] = cand[j] == c ? acc[j] + w : acc[j];
        }
        wave_sum8(acc);
        if (GW > 1) {
          if (lane == 0) {
#pragma unroll
            for (int j = 0; j < kSlots; ++j) sh_w[gw][j] = acc[j];
          }
          __syncthreads();
          if (threadIdx.x == 0) {
            for (int k = 1; k < GW; k <<= 1)
              for (int q = 0; q + k < GW; q += 2 * k)
                for (int j = 0; j < kSlots; ++j) sh_w[q][j] = sh_w[q][j] + sh_w[q + k][j];
            for (int j = 0; j < kSlots; ++j) sh_tot[j] = sh_w[0][j];
          }
          __syncthreads();
#pragma unroll
          for (int j = 0; j < kSlots; ++j) acc[j] = sh_tot[j];
        }
        argmax8(cand, acc, cb, wb);
      } else {
        uint32_t mc = kEmpty;
        float mw = 0.0f;
        for (uint64_t i = b + gl; i < e; i += P) {
          const uint32_t v = g.edg[i];
          if (v == u) continue;
          bm_step(mc, mw, s.lab[v], edge_w(g, i));
        }
        wave_bm_tree(mc, mw);
        if (GW == 1) {
          cb = __shfl(mc, 0);
          wb = __shfl(mw, 0);
        } else {
          if (lane == 0) { sh_c[gw][0] = mc; sh_w[gw][0] = mw; }
          __syncthreads();
          if (threadIdx.x == 0) {
            for (int k = 1; k < GW; k <<= 1)
              for (int q = 0; q + k < GW; q += 2 * k)
                if (sh_w[q + k][0] > sh_w[q][0]) { sh_w[q][0] = sh_w[q + k][0]; sh_c[q][0] = sh_c[q + k][0]; }
            sh_cand[0] = sh_c[0][0];
            sh_tot[0] = sh_w[0][0];
          }
          __syncthreads();
          cb = sh_cand[0];
          wb = sh_tot[0];
        }
      }
      const bool moved = wb > 0.0f && cb != d && !(pickless && cb > d);
      if (gl == 0) {
        s.aff[u] = 0;
        sa += 1;
        se += deg;
      }
      if (moved) {
        if (SYNC) {
          if (gl == 0) {
            const unsigned long long pos = atomicAdd(&s.ctr->nmoves, 1ull);
            s.moves[pos] = ((uint64_t)u << 32) | cb;
          }
        } else {
          if (gl == 0) {
            s.lab[u] = cb;
            sc += 1;
            sx += deg;
          }
          for (uint64_t i = b + gl; i < e; i += P) {
            const uint32_t v = g.edg[i];
            if (v != u) s.aff[v] = 1;
          }
        }
      }
    }
    if (GW > 1) __syncthreads();
  }
  flush_stats(s.ctr, sa, se, sc, sx);
}

// ---- synchronous mode: commit the change list ------------------------------------------------
// Lane per change for the label write; the wave then marks each change's neighbours together.
__global__ __launch_bounds__(kBlock) void k_apply(DevGraph g, DevState s) {
  if (s.ctr->done) return;
  const unsigned long long nm = s.ctr->nmoves;
  const int lane = lane_id();
  const uint64_t wave_global = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * kBlock) >> 6;
  unsigned long long sx = 0;
  for (uint64_t base = wave_global * 64; base < nm; base += nwaves * 64) {
    const uint64_t idx = base + lane;
    uint32_t u = 0, c = 0;
    uint64_t b = 0, e = 0;
    if (idx < nm) {
      const uint64_t mv = s.moves[idx];
      u = (uint32_t)(mv >> 32);
      c = (uint32_t)mv;
      s.lab[u] = c;
      b = g.off[u];
      e = g.off[(uint64_t)u + 1];
      sx += e - b;
    }
    const int cnt = (int)((nm - base) < 64 ? (nm - base) : 64);
    for (int j = 0; j < cnt; ++j) {
      const uint32_t uj = __shfl(u, j);
      const uint64_t bj = __shfl(b, j), ej = __shfl(e, j);
      for (uint64_t i = bj + lane; i < ej; i += 64) {
        const uint32_t v = g.edg[i];
        if (v != uj) s.aff[v] = 1;
      }
    }
  }
  flush_stats(s.ctr, 0, 0, 0, sx);
}

// ---- end of an iteration: convergence test on the device ------------------------------------
__global__ void k_finish(DevCounters *ctr, uint64_t n, double tolerance, int max_iterations, int pickstep,
                         int sync) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ctr->done) return;
  const unsigned int l = ctr->iter;
  const bool pickless = pickstep > 0 && (l % (unsigned int)pickstep) == 0;
  const unsigned long long C = sync ? ctr->nmoves : ctr->C;
  if (l < RLC_MAX_ITER_STATS) {
    ctr->hist[l][0] = ctr->A;
    ctr->hist[l][1] = ctr->E;
    ctr->hist[l][2] = C;
    ctr->hist[l][3] = ctr->X;
  }
  ctr->iter = l + 1;
  ctr->last_changed = C;
  // rakLowmemLoopCuU: ++l; if (!PICKLESS && double(n)/N <= E) break;  (rakLowmemCuda.hxx:486-487)
  if ((!pickless && double(C) / double(n) <= tolerance) || (int)(l + 1) >= max_iterations) ctr->done = 1;
  ctr->A = ctr->E = ctr->C = ctr->X = 0;
  ctr->nmoves = 0;
}

__global__ __launch_bounds__(kBlock) void k_init_labels(uint32_t *lab, const uint32_t *init, uint64_t n) {
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < n; u += (uint64_t)gridDim.x * kBlock)
    lab[u] = init ? init[u] : (uint32_t)u;
}

__global__ __launch_bounds__(kBlock) void k_mark_all(uint8_t *aff, uint64_t n, DevCounters *ctr, int max_iterations) {
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (uint64_t u = t; u < n; u += (uint64_t)gridDim.x * kBlock) aff[u] = 1;
  if (t == 0) {
    ctr->A = ctr->E = ctr->C = ctr->X = 0;
    ctr->nmoves = 0;
    ctr->iter = 0;
    ctr->last_changed = 0;
    ctr->done = max_iterations <= 0 ? 1u : 0u;
  }
  // per-iteration history: zeroed by the grid
  unsigned long long *h = &ctr->hist[0][0];
  for (uint64_t i = t; i < (uint64_t)RLC_MAX_ITER_STATS * 4; i += (uint64_t)gridDim.x * kBlock) h[i] = 0;
}

// ---- FP64 modularity --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mod_zero(double *ctot, uint8_t *seen, double *out, uint64_t n) {
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (uint64_t i = t; i < n; i += (uint64_t)gridDim.x * kBlock) { ctot[i] = 0.0; seen[i] = 0; }
  if (t < 4) out[t] = 0.0;
}

__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) x += __shfl_xor(x, k);
  return x;
}

// per vertex: K_u = sum of weights, intra_u = weight to same-label neighbours; tot[c] += K_u
__global__ __launch_bounds__(kBlock) void k_mod_scan(DevGraph g, const uint32_t *__restrict__ lab, double *ctot,
                                                     uint8_t *seen, double *out) {
  double si = 0.0, sk = 0.0;
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < g.n; u += (uint64_t)gridDim.x * kBlock) {
    const uint32_t c = lab[u];
    double ku = 0.0, iu = 0.0;
    for (uint64_t i = g.off[u]; i < g.off[u + 1]; ++i) {
      const double w = (double)edge_w(g, i);
      ku += w;
      if (lab[g.edg[i]] == c) iu += w;
    }
    seen[c] = 1;
    if (ku != 0.0) atomicAdd(&ctot[c], ku);
    si += iu;
    sk += ku;
  }
  si = wave_sum_f64(si);
  sk = wave_sum_f64(sk);
  if (lane_id() == 0) {
    atomicAdd(&out[0], si);
    atomicAdd(&out[1], sk);
  }
}

__global__ __launch_bounds__(kBlock) void k_mod_sum(const double *ctot, const uint8_t *seen, uint64_t n, double *out) {
  double sq = 0.0, cnt = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
    sq += ctot[i] * ctot[i];
    cnt += seen[i] ? 1.0 : 0.0;
  }
  sq = wave_sum_f64(sq);
  cnt = wave_sum_f64(cnt);
  if (lane_id() == 0) {
    atomicAdd(&out[2], sq);
    atomicAdd(&out[3], cnt);
  }
}

inline unsigned grid_for(uint64_t items, uint64_t per_block, unsigned cap) {
  uint64_t gsz = (items + per_block - 1) / per_block;
  if (gsz < 1) gsz = 1;
  if (gsz > cap) gsz = cap;
  return (unsigned)gsz;
}

constexpr unsigned kGridCap = 8192;

}  // namespace

void launch_init_labels(const DevGraph &g, const DevState &s, const uint32_t *init, hipStream_t st) {
  k_init_labels<<<grid_for(g.n, kBlock, kGridCap), kBlock, 0, st>>>(s.lab, init, g.n);
}

void launch_mark_all(const DevGraph &g, const DevState &s, int max_iterations, hipStream_t st) {
  k_mark_all<<<grid_for(g.n, kBlock, kGridCap), kBlock, 0, st>>>(s.aff, g.n, s.ctr, max_iterations);
}

template <bool SYNC, bool MG>
static void launch_moves(const DevGraph &g, const DevState &s, const ClassLists &cl, const LoopParams &p,
                         hipStream_t st) {
  const int S = cl.sweeps;
  for (int w = 0; w < S; ++w) {
    const uint64_t b0 = cl.seg[w], e0 = cl.seg[w + 1];
    const uint64_t b1 = cl.seg[S + w], e1 = cl.seg[S + w + 1];
    const uint64_t b2 = cl.seg[2 * S + w], e2 = cl.seg[2 * S + w + 1];
    if (e0 > b0)
      k_move_lane<SYNC, MG><<<grid_for(e0 - b0, kBlock, kGridCap), kBlock, 0, st>>>(g, s, cl.list + b0, e0 - b0,
                                                                                    p.pickstep);
    if (e1 > b1)
      k_move_group<SYNC, MG, 64><<<grid_for(e1 - b1, kBlock / 64, kGridCap), kBlock, 0, st>>>(
          g, s, cl.list + b1, e1 - b1, p.pickstep);
    if (e2 > b2)
      k_move_group<SYNC, MG, 256><<<grid_for(e2 - b2, 1, kGridCap), kBlock, 0, st>>>(g, s, cl.list + b2, e2 - b2,
                                                                                     p.pickstep);
  }
}

void launch_iteration(const DevGraph &g, const DevState &s, const ClassLists &cl, const LoopParams &p,
                      hipStream_t st, hipEvent_t ev_move0, hipEvent_t ev_move1) {
  if (ev_move0) (void)hipEventRecord(ev_move0, st);
  if (p.sync) {
    if (p.mg) launch_moves<true, true>(g, s, cl, p, st);
    else launch_moves<true, false>(g, s, cl, p, st);
  } else {
    if (p.mg) launch_moves<false, true>(g, s, cl, p, st);
    else launch_moves<false, false>(g, s, cl, p, st);
  }
  if (ev_move1) (void)hipEventRecord(ev_move1, st);
  if (p.sync) k_apply<<<grid_for(g.n, kBlock, 4096), kBlock, 0, st>>>(g, s);
  k_finish<<<1, 64, 0, st>>>(s.ctr, g.n, p.tolerance, p.max_iterations, p.pickstep, p.sync ? 1 : 0);
}

void launch_modularity(const DevGraph &g, const uint32_t *lab, double *ctot, uint8_t *seen, double *out,
                       hipStream_t st) {
  k_mod_zero<<<grid_for(g.n, kBlock, kGridCap), kBlock, 0, st>>>(ctot, seen, out, g.n);
  k_mod_scan<<<grid_for(g.n, kBlock, kGridCap), kBlock, 0, st>>>(g, lab, ctot, seen, out);
  k_mod_sum<<<grid_for(g.n, kBlock, kGridCap), kBlock, 0, st>>>(ctot, seen, g.n, out);
}

}  // namespace rlc
