// Memory helpers of the multi-lane chunk kernels (group reruns, DPP summaries):
// float4 row pieces that may or may not be 16-byte aligned, and the barriers
// that keep a prefetch ring a prefetch.
#pragma once

#include "tlpc_common.h"

namespace tlpc {
namespace vmem {

// A float4 at p: one 16-byte access when aligned (AL), four 4-byte ones if not
// (the packed coefficient rows of a ragged batch start wherever the signals
// before them end).
template <bool AL>
__device__ __forceinline__ float4 ldq(const float* p) {
  if constexpr (AL) {
    return *reinterpret_cast<const float4*>(p);
  } else {
    return make_float4(p[0], p[1], p[2], p[3]);
  }
}
template <bool AL>
__device__ __forceinline__ void stq(float* p, float4 v) {
  if constexpr (AL) {
    *reinterpret_cast<float4*>(p) = v;
  } else {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
    p[3] = v.w;
  }
}

// Keeps a step's memory instructions where the source puts them: without it the
// scheduler may gather a block's ring reloads right before the next block
// consumes them, and the prefetch distance drops to nothing.
__device__ __forceinline__ void pin() { __builtin_amdgcn_sched_barrier(0); }

// Ring fills pass through an empty asm statement: a ring slot is a loop phi of
// two loads (the prologue's and the previous iteration's), which InstCombine
// would otherwise fold into one load at its use.
__device__ __forceinline__ float opq(float v) {
  asm("" : "+v"(v));
  return v;
}
__device__ __forceinline__ float2 opq(float2 v) {
  asm("" : "+v"(v.x), "+v"(v.y));
  return v;
}
__device__ __forceinline__ float4 opq(float4 v) {
  asm("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
  return v;
}

// A loop-carried state read from memory waits for its loads here, once, and not
// at its first use inside the loop, where the wait-count pass would repeat the
// wait (and drain the ring loads) on every iteration.
template <int K>
__device__ __forceinline__ void settle(double (&v)[K]) {
#pragma unroll
  for (int i = 0; i < K; ++i) asm volatile("" : "+v"(v[i]));
}

}  // namespace vmem
}  // namespace tlpc
