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
  // Ragged batch (null: a uniform [B, T] batch). B signals of one order n packed
  // back to back; rg is a device table of int64:
  //   rg[0..B)          sample offset of each signal in x / y / g / dx
  //   rg[B..2B)         element offset of A[0, 0] (and dA) of each signal ([T_k, n] rows)
  //   rg[2B..3B)        element offset of zi / dzi of each signal
  //   rg[3B..4B)        length T_k
  //   rg[4B..5B]        first chunk id of each signal (rg[5B] = G, the chunk count)
  //   rg[5B+1..5B+1+G)  signal of each chunk
  // Summaries and carries are indexed by chunk id; P holds the largest P_k.
  const int64_t* rg;
  int64_t G;
  // Time-invariant adjoint with da partials (float4 streams): per signal, the y
  // window's values below position 0 (y_{-1-k} = zi[k], zero past the order)
  // as ti_ypad(NP) floats ending at position -1, 16-byte aligned; or null.
  S* ypad;
  // Long carry chains of a uniform batch at low orders: nseg segments per signal
  // (0: one workgroup walks each chain), with their maps and incoming states in
  // segs (workspace).
  typename Wide<S>::type* segs;
  int64_t nseg;
  // Adjoint without a saved y (null: y is read): the forward's input and its
  // chunk carries [B, P, NP], from which the adjoint rerun rebuilds y.
  const S* rx;
  const typename Wide<S>::type* fcarry;
  // Adjoint of a ragged batch whose summary slots still hold the forward's
  // transition matrices (slot of chunk c: Phi_c): the adjoint chain reads them
  // transposed (Psi_c = Phi_c^T) and only the zero-state responses are formed.
  int kept;
};

// Floats per signal of ChunkArgs::ypad: the y window of an order-NP adjoint
// reaches position -1-NP; whole 4-sample blocks.
constexpr int ti_ypad(int np) { return (np + 1 + 3) / 4 * 4; }

// One signal's view of the batch (uniform or ragged).
struct Span {
  int64_t T;      // length
  int64_t P;      // chunks
  int64_t xo;     // offset of sample 0 in x / y / g / dx
  int64_t ao;     // offset of A[0, 0]
  int64_t at;     // element stride of A over time
  int64_t dao;    // offset of dA[0, 0] (per-sample gradient)
  int64_t zo;     // offset of zi / dzi
  int64_t cbase;  // chunk id of chunk 0 (carry index = cbase + c)
};

template <typename S>
__device__ __forceinline__ Span sig_span(const ChunkArgs<S>& a, int64_t k) {
  Span s;
  if (a.rg) {
    const int64_t K = a.B;
    s.xo = a.rg[k];
    s.ao = a.rg[K + k];
    s.zo = a.rg[2 * K + k];
    s.T = a.rg[3 * K + k];
    s.cbase = a.rg[4 * K + k];
    s.P = a.rg[4 * K + k + 1] - s.cbase;
    s.at = a.n;
    s.dao = s.ao;
  } else {
    s.xo = k * a.T;
    s.ao = k * a.a_bstride;
    s.zo = k * a.n;
    s.T = a.T;
    s.cbase = k * a.P;
    s.P = a.P;
    s.at = a.a_tstride;
    s.dao = k * a.T * a.n;
  }
  return s;
}

// Signal of chunk id `cid` (carry index).
template <typename S>
__device__ __forceinline__ int64_t chunk_signal(const ChunkArgs<S>& a, int64_t cid) {
  return a.rg ? a.rg[5 * a.B + 1 + cid] : cid / a.P;
}

// Summary slot of chunk c of a signal. Uniform batches store the forward
// summaries of chunks 0..P-2 and the adjoint ones of chunks 1..P-1 densely;
// ragged batches index them by chunk id.
template <typename S>
__device__ __forceinline__ int64_t summ_slot(const ChunkArgs<S>& a, int64_t k, const Span& s,
                                             int64_t c, bool fwd) {
  if (a.rg) return s.cbase + c;
  return k * (a.P - 1) + (fwd ? c : c - 1);
}

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

// Reruns with NP / 4 lanes per chunk (float32 time-varying filters of order 8,
// 16 and 32 whose rows are exactly NP wide, uniform or ragged; not the adjoint
// that rebuilds y, nor time-invariant da partials).
bool grp_rerun_ok(const ChunkArgs<float>& a, int np);
int launch_grp_fwd_rerun(ChunkArgs<float> a, int np, hipStream_t st);
int launch_grp_bwd_rerun(ChunkArgs<float> a, int np, hipStream_t st);
// Adjoint zero-state responses only, for a ragged backward handed the forward's
// summaries (ChunkArgs::kept): float32 orders 8, 16 and 32.
bool grp_wsum_ok(const ChunkArgs<float>& a, int np);
int launch_grp_wsum(ChunkArgs<float> a, int np, hipStream_t st);
// The summary work item of a chunk's lanes in the multi-lane float32 summary
// kernels, numbered as the LDS-staged summaries number them: item i of a
// uniform batch is chunk (i / (P-1), i % (P-1)) (+1 for the adjoint); in a
// ragged batch item i is chunk id i, and the chunks without a summary (a
// signal's last forward, its first adjoint) are idle.
struct SumItem {
  const float* Ab;  // A at the chunk's first sample
  const float* xb;  // x (forward) or g (adjoint) at the chunk's first sample
  double* out;
  int live;  // samples of the signal in the chunk
};

template <int NP, bool FWD>
__device__ __forceinline__ bool summary_item(const ChunkArgs<float>& a, int64_t item, SumItem& it) {
  int64_t k, c;
  if (a.rg) {
    if (item >= a.G) return false;
    k = chunk_signal(a, item);
    c = item - sig_span(a, k).cbase;
  } else {
    const int64_t cps = a.P - 1;
    if (item >= a.B * cps) return false;
    k = item / cps;
    c = item - k * cps + (FWD ? 0 : 1);
  }
  const Span sp = sig_span(a, k);
  if (FWD ? (c >= sp.P - 1) : (c < 1)) return false;
  const int64_t t = c * a.L;
  const int64_t rem = sp.T - t;
  it.Ab = a.A + sp.ao + t * NP;
  it.xb = a.x + sp.xo + t;
  it.out = a.summ + item * (NP * NP + NP);
  it.live = rem < a.L ? (int)rem : a.L;
  return true;
}

// Multi-lane chunk summaries: float32 time-varying orders 16 and 32 on DPP
// row broadcasts (tlpc_row_summary.hip), 2, 4 and 8 on quads (tlpc_quad_rerun.hip);
// n == NP, plain rows, L a multiple of max(NP, 16).
bool row_summary_ok(const ChunkArgs<float>& a, int np);
int launch_row_summary(ChunkArgs<float> a, int np, bool fwd, hipStream_t st);

// Per-signal tiled kernels (order 16, float32, time-varying): one workgroup per
// signal, no workspace. a.y may be null in the backward when dA is not wanted.
int launch_tile16_fwd(ChunkArgs<float> a, hipStream_t st);
int launch_tile16_bwd(ChunkArgs<float> a, hipStream_t st);

// Per-signal time-invariant kernels (orders 16..32, float32, coefficients a[B, n]):
// one workgroup per signal, block products on MFMA, float64 state transitions,
// no workspace. The backward takes the gradient in a.x, writes a.dx, a.dzi and
// the reduced a.da_ti (each optional), and reads a.y only for da.
bool ti_mfma_ok(int n, int64_t T);
int launch_ti_mfma_fwd(ChunkArgs<float> a, hipStream_t st);
int launch_ti_mfma_bwd(ChunkArgs<float> a, hipStream_t st);

// Coefficient as used by the recurrence: A, or -w for the first-order scan.
template <typename S>
__device__ __forceinline__ S coef(const ChunkArgs<S>& a, S v) {
  return a.neg ? -v : v;
}

// Staging block of the chunked kernels for a compiled order (chunk lengths are
// multiples of it).
constexpr int chunk_block(int np) { return np < 16 ? 16 : np; }

}  // namespace tlpc
