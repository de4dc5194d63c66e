// Memory helpers of the multi-lane chunk kernels (group and quad reruns, DPP
// and quad summaries): float4 row pieces that may or may not be 16-byte
// aligned, and the barriers that keep a prefetch ring a prefetch.
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

// y_u of one signal for the adjoint's window (y_{-1-k} = zi[k] for k < n, zero
// past the order and at or past T): an always-valid address, and whether its
// value counts. Both are formed without a branch: a branch in the step loop
// lets the wait-count pass drain the prefetch ring at its join, and so does a
// select on a value right after its load, so the ring keeps the raw value and
// the step that consumes it applies the select. zb is the signal's zi, or any
// valid address when there is none (haszi false).
// Both candidate offsets pass through an empty asm statement before the select
// (on offsets, not pointers, which would lose their global address space), and
// the flag is formed bitwise: otherwise the compiler turns the conditionals into
// control flow.
__device__ __forceinline__ const float* yw_ptr(const float* yb, const float* zb, int n, int64_t T, int64_t u) {
  const int64_t k = -u - 1;
  int64_t op = u < T ? u : T - 1;
  int64_t on = k < n ? k : n - 1;
  asm("" : "+v"(op), "+v"(on));
  const bool pos = u >= 0;
  return (pos ? yb : zb) + (pos ? op : on);
}
__device__ __forceinline__ bool yw_flag(bool haszi, int n, int64_t T, int64_t u) {
  const int pos = u >= 0;
  int ok = (pos & (int)(u < T)) | ((pos ^ 1) & (int)haszi & (int)(-u - 1 < n));
  asm("" : "+v"(ok));
  return ok != 0;
}

}  // namespace vmem
}  // namespace tlpc
