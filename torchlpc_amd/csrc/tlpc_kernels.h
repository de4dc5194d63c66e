// Internal interface between the C ABI (tlpc_api.hip) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tlpc_common.h"

namespace tlpc {

// Arguments shared by the chunked, per-signal and serial strategies. For the
// adjoint, `x` carries the upstream gradient g and `y` the saved forward output.
template <typename S>
struct ChunkArgs {
  const S* x;
  const S* A;
  const S* zi;
  S* y;
  S* dx;
  S* dA;      // time-varying dA [B,T,n] (adjoint) or null
  S* dzi;     // [B,n] or null
  typename Wide<S>::type* dapart;  // time-invariant da: per-chunk partials [B,P,NP] (workspace) or null
  S* da_ti;   // time-invariant da output [B,n] or null
  int64_t B, T;
  int n;      // filter order
  int L;      // chunk length (multiple of the staging block)
  int64_t P;  // chunks per signal
  int64_t a_bstride, a_tstride;  // element strides of A over signals and time (0: time-invariant)
  int neg;    // 1: the coefficient array holds -A (first-order scan weights w = -A)
  typename Wide<S>::type* summ;   // chunk summaries [B, P-1, NP*NP+NP]
  typename Wide<S>::type* carry;  // chunk carries [B, P, NP]
  typename Wide<S>::type* wsum;   // order-16 adjoint: zero-state responses [B, P-1, NP]
};

template <typename S, int NP>
int launch_chunked_fwd(ChunkArgs<S> a, hipStream_t st);
template <typename S, int NP>
int launch_chunked_bwd(ChunkArgs<S> a, hipStream_t st);

template <typename S>
int launch_serial_fwd(ChunkArgs<S> a, hipStream_t st);
template <typename S>
int launch_serial_bwd(ChunkArgs<S> a, hipStream_t st);

template <typename S>
int launch_fill_zero(S* p, int64_t count, hipStream_t st);

// Order-16 float32 chunked-scan kernels (DPP matrix rows, quad reruns).
// Forward summaries: transition matrices Phi and zero-state responses z; with
// with_z false only Phi (an adjoint that has no forward summaries to reuse).
int launch_chunk16_fwd_summary(ChunkArgs<float> a, bool with_z, hipStream_t st);
// Adjoint zero-state responses w (into a.wsum); the adjoint transition matrices
// are the transposes of the forward Phi.
int launch_chunk16_adj_summary(ChunkArgs<float> a, hipStream_t st);
int launch_chunk16_carry(ChunkArgs<float> a, bool fwd, hipStream_t st);
int launch_chunk16_fwd_rerun(ChunkArgs<float> a, hipStream_t st);
int launch_chunk16_adj_rerun(ChunkArgs<float> a, hipStream_t st);

// Per-signal fused kernels (order 16, float32): one workgroup per signal.
int launch_signal_fwd16(ChunkArgs<float> a, hipStream_t st);
int signal_chunk_len(int64_t T, int requested);

// Coefficient as used by the recurrence: A, or -w for the first-order scan.
template <typename S>
__device__ __forceinline__ S coef(const ChunkArgs<S>& a, S v) {
  return a.neg ? -v : v;
}

// Staging block of the chunked kernels for a compiled order (chunk lengths are
// multiples of it).
constexpr int chunk_block(int np) { return np < 16 ? 16 : np; }

}  // namespace tlpc
