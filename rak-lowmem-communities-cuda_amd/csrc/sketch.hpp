// Device sketch primitives for gfx950 (wave64). A sketch lives in one lane's VGPRs: 8 labels and
// 8 float weights for Misra-Gries, one (label, weight) pair for Boyer-Moore. Group-wide merges
// and sums run over lanes with cross-lane moves in a fixed binary tree, so that every result
// is a deterministic function of the edge order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rlc {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kSlots = 8;

// Weighted Misra-Gries sketch with the slot policy of the reference CPU path
// (inc/rakLowmem.hxx:63-89): add w to the slot holding c; otherwise take the highest-index
// zero-weight slot; otherwise subtract w from every slot, clamped at zero. (The reference GPU
// variant, inc/sketchCuda.hxx:73-108, takes the lowest free slot and does not clamp.)
struct MG8 {
  uint32_t c[kSlots];
  float w[kSlots];

  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int i = 0; i < kSlots; ++i) { c[i] = kEmpty; w[i] = 0.0f; }
  }

  __device__ __forceinline__ void accumulate(uint32_t lab, float wt) {
    bool has = false;
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
      const bool m = c[i] == lab;
      w[i] = m ? w[i] + wt : w[i];
      has = has || m;
    }
    if (has) return;
    int f = -1;
#pragma unroll
    for (int i = 0; i < kSlots; ++i) f = (w[i] == 0.0f) ? i : f;
    if (f >= 0) {
#pragma unroll
      for (int i = 0; i < kSlots; ++i)
        if (i == f) { c[i] = lab; w[i] = wt; }
    } else {
#pragma unroll
      for (int i = 0; i < kSlots; ++i) {
        const float x = w[i] - wt;
        w[i] = x < 0.0f ? 0.0f : x;
      }
    }
  }

  // Merge: accumulate each non-zero slot of o, in slot order (inc/sketchCuda.hxx:150-158).
  __device__ __forceinline__ void absorb(const uint32_t (&oc)[kSlots], const float (&ow)[kSlots]) {
#pragma unroll
    for (int j = 0; j < kSlots; ++j)
      if (ow[j] != 0.0f) accumulate(oc[j], ow[j]);
  }
};

// First slot of strictly largest weight, from (label 0, weight 0) (inc/rakLowmem.hxx:164-171).
__device__ __forceinline__ void argmax8(const uint32_t (&c)[kSlots], const float (&w)[kSlots], uint32_t &cb,
                                        float &wb) {
  cb = 0;
  wb = 0.0f;
#pragma unroll
  for (int i = 0; i < kSlots; ++i)
    if (w[i] > wb) { wb = w[i]; cb = c[i]; }
}

// Weighted Boyer-Moore majority step (inc/rakLowmem.hxx:121-125, inc/rakLowmemCuda.hxx:49-51).
__device__ __forceinline__ void bm_step(uint32_t &mc, float &mw, uint32_t c, float w) {
  if (c == mc) mw += w;
  else if (mw > w) mw -= w;
  else { mc = c; mw = w; }
}

// ---- wave-level helpers ---------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ uint32_t uniform_u32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) x += __shfl_xor(x, k);
  return x;
}

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Merge tree over the 64 lanes of a wave: at level k (1, 2, 4, .., 32) lane p with p % 2k == 0
// absorbs lane p+k. The result is in lane 0; other lanes hold partial sketches.
__device__ __forceinline__ void wave_merge_tree(MG8 &s) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 1; k < 64; k <<= 1) {
    uint32_t oc[kSlots];
    float ow[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      oc[j] = __shfl_down(s.c[j], k);
      ow[j] = __shfl_down(s.w[j], k);
    }
    if ((lane & (2 * k - 1)) == 0) s.absorb(oc, ow);
  }
}

// Boyer-Moore max tree over the wave: lane p takes lane p+k's candidate if strictly heavier.
__device__ __forceinline__ void wave_bm_tree(uint32_t &mc, float &mw) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 1; k < 64; k <<= 1) {
    const uint32_t oc = __shfl_down(mc, k);
    const float ow = __shfl_down(mw, k);
    if ((lane & (2 * k - 1)) == 0 && ow > mw) { mc = oc; mw = ow; }
  }
}

// Butterfly sum over the wave with masks 1, 2, .., 32: every lane ends with the value of the
// binary tree (p absorbs p+k) evaluated at its group's root; float addition is commutative, so
// all lanes agree bit for bit.
__device__ __forceinline__ void wave_sum8(float (&a)[kSlots]) {
#pragma unroll
  for (int k = 1; k < 64; k <<= 1)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) a[j] = a[j] + __shfl_xor(a[j], k);
}

}  // namespace rlc
