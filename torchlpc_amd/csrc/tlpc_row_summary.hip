// Chunk summaries (phase 1 of the chunked scan) on DPP row broadcasts, for
// float32 time-varying filters of order 16 and 32 in uniform or ragged batches,
// NP lanes per chunk. A 16-lane DPP row loads four consecutive coefficient rows
// (4 * NP floats) as NP / 16 float4 per lane, converts them to float64, and
// every lane takes coefficient k of a step through a row_newbcast operand of
// v_fmac_f64: no LDS staging. (Read from LDS, an order-32 row of float64
// coefficients costs 16 ds_read_b128 per wave and step against 32 FMAs, which
// left the LDS-staged summaries bound by LDS.)
//
//   forward  chunks 0 .. P-2: lane k sweeps the chunk backwards with the
//            transposed recurrence from a unit at output k (sample L-1-k): the
//            pending sums left before the chunk are row k of Phi and the
//            lambdas met on the way, dotted with x, give z[k].
//   adjoint  chunks 1 .. P-1: lane r runs the zero-input recurrence forwards
//            from a unit at the state entering the chunk (position -1-r): its
//            final window is row r of Psi, and w[r] = sum_t g_t h_r(t).
//
// Layouts as in the LDS-staged summaries: summ[item][k * NP + r] = Phi[k][r]
// (Psi[k][r] for the adjoint) and summ[item][NP * NP + k] = z[k] (w[k]).
// Arithmetic is float64 on float32 data. Reference recurrence:
// torchlpc/csrc/scan_cpu.cpp:112-120; adjoint: torchlpc/core.py:186-229.
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"
#include "tlpc_vmem.h"

namespace tlpc {
namespace rsum {

using namespace vmem;
using dpp16::fma_bcast;

// 64-lane workgroups (two chunks at order 32, four at order 16): a launch of a
// few thousand chunks then spreads evenly over the CUs.
constexpr int THREADS = 64;

// Where coefficient k of row jr of a 4-row group sits: lane L0(jr) + (k >> 2) of
// the 16-lane row, component k & 3 of the lane's float4 number D(jr).
//   order 16: lane l holds row l >> 2, coefficients 4 (l & 3) .. 4 (l & 3) + 3
//             (L0 = 4 jr, D = 0);
//   order 32: lane l holds, as float4 m, row (l >> 3) + 2m, coefficients
//             4 (l & 7) .. 4 (l & 7) + 3 (L0 = 8 (jr & 1), D = jr >> 1).
// Either way the row loads a group as one contiguous 16 * NP-byte piece.
template <int NP>
struct Lay {
  static constexpr int NF = NP / 16;  // float4 per lane per 4-row group
  static constexpr int NG = NP / 4;   // 4-row groups per block of NP steps
  __device__ static int row(int l, int m) { return NP == 16 ? (l >> 2) : (l >> 3) + 2 * m; }
  __device__ static int col(int l) { return NP == 16 ? 4 * (l & 3) : 4 * (l & 7); }
  template <int JR>
  static constexpr int L0 = NP == 16 ? 4 * JR : 8 * (JR & 1);
  template <int JR>
  static constexpr int D = NP == 16 ? 0 : (JR >> 1);
};

// 16 products of the transposed step: %0..%15 -= coefficient (component
// %16..%19 of lane %21..%24) * lambda (%20).
#define TLPC_RS_ADJ16 \
  "v_fmac_f64_dpp %0, -%16, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %1, -%17, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %2, -%18, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %3, -%19, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %4, -%16, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %5, -%17, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %6, -%18, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %7, -%19, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %8, -%16, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %9, -%17, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %10, -%18, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %11, -%19, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %12, -%16, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %13, -%17, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %14, -%18, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %15, -%19, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t"

// 16 products of the forward step, oldest lag first: accumulator %(f & 3) -=
// coefficient (component %(20 + (f & 3)) of lane %(24 + (f >> 2))) * window
// value %(4 + f), for f = 15 .. 0.
#define TLPC_RS_FF16 \
  "v_fmac_f64_dpp %3, -%23, %19 row_newbcast:%27 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %2, -%22, %18 row_newbcast:%27 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %1, -%21, %17 row_newbcast:%27 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %0, -%20, %16 row_newbcast:%27 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %3, -%23, %15 row_newbcast:%26 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %2, -%22, %14 row_newbcast:%26 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %1, -%21, %13 row_newbcast:%26 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %0, -%20, %12 row_newbcast:%26 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %3, -%23, %11 row_newbcast:%25 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %2, -%22, %10 row_newbcast:%25 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %1, -%21, %9 row_newbcast:%25 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %0, -%20, %8 row_newbcast:%25 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %3, -%23, %7 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %2, -%22, %6 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %1, -%21, %5 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t" \
  "v_fmac_f64_dpp %0, -%20, %4 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t"

// Transposed step at slot J = t mod NP of the pending-sum window p (p[u mod NP]
// is the pending sum of position u): lambda = p[J]; p[(J - i) mod NP] -=
// A[t, i] * lambda for i = 1..NP, where i = NP lands on slot J itself, which now
// stands for position t - NP and starts from zero (`fresh`). Returns lambda.
// The products are independent of each other; the next step waits only for the
// first. The s_nop covers a conversion right before (a DPP source needs two
// wait states).
template <int NP, int J>
__device__ __forceinline__ double adj_step(double (&p)[NP], const double (&d)[NP / 16][4]) {
  constexpr int L0 = Lay<NP>::template L0<(J & 3)>;
  const double(&e)[4] = d[Lay<NP>::template D<(J & 3)>];
  const double lam = p[J];
  double fresh = 0.0;
  if constexpr (NP == 16) {
    asm("s_nop 1\n\t" TLPC_RS_ADJ16
        : "+v"(p[(J - 1) & 15]), "+v"(p[(J - 2) & 15]), "+v"(p[(J - 3) & 15]), "+v"(p[(J - 4) & 15]),
          "+v"(p[(J - 5) & 15]), "+v"(p[(J - 6) & 15]), "+v"(p[(J - 7) & 15]), "+v"(p[(J - 8) & 15]),
          "+v"(p[(J - 9) & 15]), "+v"(p[(J - 10) & 15]), "+v"(p[(J - 11) & 15]), "+v"(p[(J - 12) & 15]),
          "+v"(p[(J - 13) & 15]), "+v"(p[(J - 14) & 15]), "+v"(p[(J - 15) & 15]), "+v"(fresh)
        : "v"(e[0]), "v"(e[1]), "v"(e[2]), "v"(e[3]), "v"(lam), "i"(L0), "i"(L0 + 1), "i"(L0 + 2), "i"(L0 + 3));
  } else {
    asm("s_nop 1\n\t" TLPC_RS_ADJ16
        : "+v"(p[(J - 1) & 31]), "+v"(p[(J - 2) & 31]), "+v"(p[(J - 3) & 31]), "+v"(p[(J - 4) & 31]),
          "+v"(p[(J - 5) & 31]), "+v"(p[(J - 6) & 31]), "+v"(p[(J - 7) & 31]), "+v"(p[(J - 8) & 31]),
          "+v"(p[(J - 9) & 31]), "+v"(p[(J - 10) & 31]), "+v"(p[(J - 11) & 31]), "+v"(p[(J - 12) & 31]),
          "+v"(p[(J - 13) & 31]), "+v"(p[(J - 14) & 31]), "+v"(p[(J - 15) & 31]), "+v"(p[(J - 16) & 31])
        : "v"(e[0]), "v"(e[1]), "v"(e[2]), "v"(e[3]), "v"(lam), "i"(L0), "i"(L0 + 1), "i"(L0 + 2), "i"(L0 + 3));
    asm("s_nop 1\n\t" TLPC_RS_ADJ16
        : "+v"(p[(J - 17) & 31]), "+v"(p[(J - 18) & 31]), "+v"(p[(J - 19) & 31]), "+v"(p[(J - 20) & 31]),
          "+v"(p[(J - 21) & 31]), "+v"(p[(J - 22) & 31]), "+v"(p[(J - 23) & 31]), "+v"(p[(J - 24) & 31]),
          "+v"(p[(J - 25) & 31]), "+v"(p[(J - 26) & 31]), "+v"(p[(J - 27) & 31]), "+v"(p[(J - 28) & 31]),
          "+v"(p[(J - 29) & 31]), "+v"(p[(J - 30) & 31]), "+v"(p[(J - 31) & 31]), "+v"(fresh)
        : "v"(e[0]), "v"(e[1]), "v"(e[2]), "v"(e[3]), "v"(lam), "i"(L0 + 4), "i"(L0 + 5), "i"(L0 + 6),
          "i"(L0 + 7));
  }
  p[J] = fresh;
  return lam;
}

// Forward step of the zero-input recurrence at slot J = t mod NP of the window h
// (h[u mod NP] = h(u)): h(t) = -sum_{i=1..NP} A[t, i] h(t - i), in four
// accumulators with the newest term (i = 1) last, so that a step waits for the
// previous one only at its last product. Returns h(t).
template <int NP, int J>
__device__ __forceinline__ double ff_step(double (&h)[NP], const double (&d)[NP / 16][4]) {
  constexpr int L0 = Lay<NP>::template L0<(J & 3)>;
  const double(&e)[4] = d[Lay<NP>::template D<(J & 3)>];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if constexpr (NP == 32) {
    asm("s_nop 1\n\t" TLPC_RS_FF16
        : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
        : "v"(h[(J - 17) & 31]), "v"(h[(J - 18) & 31]), "v"(h[(J - 19) & 31]), "v"(h[(J - 20) & 31]),
          "v"(h[(J - 21) & 31]), "v"(h[(J - 22) & 31]), "v"(h[(J - 23) & 31]), "v"(h[(J - 24) & 31]),
          "v"(h[(J - 25) & 31]), "v"(h[(J - 26) & 31]), "v"(h[(J - 27) & 31]), "v"(h[(J - 28) & 31]),
          "v"(h[(J - 29) & 31]), "v"(h[(J - 30) & 31]), "v"(h[(J - 31) & 31]), "v"(h[(J - 32) & 31]),
          "v"(e[0]), "v"(e[1]), "v"(e[2]), "v"(e[3]), "i"(L0 + 4), "i"(L0 + 5), "i"(L0 + 6), "i"(L0 + 7));
  }
  asm("s_nop 1\n\t" TLPC_RS_FF16
      : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
      : "v"(h[(J - 1) & (NP - 1)]), "v"(h[(J - 2) & (NP - 1)]), "v"(h[(J - 3) & (NP - 1)]),
        "v"(h[(J - 4) & (NP - 1)]), "v"(h[(J - 5) & (NP - 1)]), "v"(h[(J - 6) & (NP - 1)]),
        "v"(h[(J - 7) & (NP - 1)]), "v"(h[(J - 8) & (NP - 1)]), "v"(h[(J - 9) & (NP - 1)]),
        "v"(h[(J - 10) & (NP - 1)]), "v"(h[(J - 11) & (NP - 1)]), "v"(h[(J - 12) & (NP - 1)]),
        "v"(h[(J - 13) & (NP - 1)]), "v"(h[(J - 14) & (NP - 1)]), "v"(h[(J - 15) & (NP - 1)]),
        "v"(h[(J - 16) & (NP - 1)]), "v"(e[0]), "v"(e[1]), "v"(e[2]), "v"(e[3]), "i"(L0), "i"(L0 + 1),
        "i"(L0 + 2), "i"(L0 + 3));
  const double y = a0 + (a1 + (a2 + a3));
  h[J] = y;
  return y;
}

#undef TLPC_RS_ADJ16
#undef TLPC_RS_FF16

template <int NP>
__device__ __forceinline__ void to_f64(double (&d)[NP / 16][4], const float4 (&c)[NP / 16]) {
#pragma unroll
  for (int m = 0; m < NP / 16; ++m) {
    d[m][0] = (double)c[m].x;
    d[m][1] = (double)c[m].y;
    d[m][2] = (double)c[m].z;
    d[m][3] = (double)c[m].w;
  }
}

// ---------------------------------------------------------------------------
// Forward summary. A block of NP steps is 4-row groups NG-1 .. 0, walked down;
// the ring holds four groups (16 steps) per lane: slot G & 3 is converted, then
// reloaded with the group 16 steps below (in this block, or in the block below;
// block 0 again after the last). Forward summaries cover whole chunks only, so
// every row read is in the chunk. x of step J of a block sits in lane J & 15 of
// xv[J >> 4].
template <int NP, int J, int JEND>
__device__ __forceinline__ void fwd_steps(double (&p)[NP], double& z, const double (&d)[NP / 16][4],
                                          const double (&xv)[NP / 16]) {
  if constexpr (J >= JEND) {
    const double lam = adj_step<NP, J>(p, d);
    fma_bcast<(J & 15)>(z, xv[J >> 4], lam);
    fwd_steps<NP, J - 1, JEND>(p, z, d, xv);
  }
}

template <int NP, bool AL, int G>
__device__ __forceinline__ void fwd_block(double (&p)[NP], double& z, float4 (&cq)[4][NP / 16],
                                          const double (&xv)[NP / 16], const float* lp, int jb, int jn) {
  if constexpr (G >= 0) {
    using Y = Lay<NP>;
    constexpr int S = G & 3;
    double d[Y::NF][4];
    to_f64<NP>(d, cq[S]);
    const int jr = G >= 4 ? jb + 4 * (G - 4) : jn + 4 * (G + Y::NG - 4);
#pragma unroll
    for (int m = 0; m < Y::NF; ++m) cq[S][m] = ldq<AL>(lp + (jr + 2 * m) * NP);
    pin();
    fwd_steps<NP, 4 * G + 3, 4 * G>(p, z, d, xv);
    pin();
    fwd_block<NP, AL, G - 1>(p, z, cq, xv, lp, jb, jn);
  }
}

template <int NP, bool AL>
__device__ __forceinline__ void fwd_chunk(const SumItem& it, int L, int k) {
  using Y = Lay<NP>;
  constexpr int NS = NP / 16;  // 16-step x pieces per block
  const int l = k & 15;
  // the lane's float4 m of the group at row jq: A[jq + row(l, m), col(l) ..];
  // row(l, m) = row(l, 0) + 2m
  const float* lp = it.Ab + Y::row(l, 0) * NP + Y::col(l);
  double p[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) p[i] = (i == NP - 1 - k) ? 1.0 : 0.0;  // unit at sample L-1-k
  double z = 0.0;
  float4 cq[4][Y::NF];
  int jb = L - NP;
#pragma unroll
  for (int g = Y::NG - 4; g < Y::NG; ++g)
#pragma unroll
    for (int m = 0; m < Y::NF; ++m) cq[g & 3][m] = opq(ldq<AL>(lp + (jb + 4 * g + 2 * m) * NP));
  float xs[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) xs[s] = opq(it.xb[jb + 16 * s + l]);
  for (; jb >= 0; jb -= NP) {
    const int jn = jb >= NP ? jb - NP : 0;
    double xv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      xv[s] = (double)xs[s];
      xs[s] = it.xb[jn + 16 * s + l];
    }
    fwd_block<NP, AL, Y::NG - 1>(p, z, cq, xv, lp, jb, jn);
  }
  // the pending sum of position -1-r (before the chunk) sits in slot NP-1-r
#pragma unroll
  for (int r = 0; r < NP; ++r) it.out[k * NP + r] = p[NP - 1 - r];
  it.out[NP * NP + k] = z;
}

template <int NP>
__global__ __launch_bounds__(THREADS) void row_fwd_summary(ChunkArgs<float> a) {
  const int64_t item = ((int64_t)blockIdx.x * THREADS + threadIdx.x) / NP;
  const int k = (int)(threadIdx.x & (NP - 1));
  SumItem it;
  if (!summary_item<NP, true>(a, item, it)) return;  // a chunk's lanes leave together
  if ((reinterpret_cast<uintptr_t>(it.Ab) & 15) == 0)
    fwd_chunk<NP, true>(it, a.L, k);
  else
    fwd_chunk<NP, false>(it, a.L, k);
}

// ---------------------------------------------------------------------------
// Adjoint summary, ascending: 4-row groups 0 .. NG-1 of a block; slot G & 3 is
// converted, then reloaded with the group 16 steps above (the block above; the
// last block again at the end). The last chunk of a signal may be partial: rows
// at or past `live` load at a clamped row and count as zero, and so do their g,
// so they add nothing to w (Psi of a signal's last chunk meets the zero state).
template <int NP, int J, int JEND>
__device__ __forceinline__ void adj_steps(double (&h)[NP], double& w, const double (&d)[NP / 16][4],
                                          const double (&gv)[NP / 16]) {
  if constexpr (J <= JEND) {
    const double y = ff_step<NP, J>(h, d);
    fma_bcast<(J & 15)>(w, gv[J >> 4], y);
    adj_steps<NP, J + 1, JEND>(h, w, d, gv);
  }
}

template <int NP, bool AL, int G>
__device__ __forceinline__ void adj_block(double (&h)[NP], double& w, float4 (&cq)[4][NP / 16],
                                          const double (&gv)[NP / 16], const float* cp, int r0, int live,
                                          int jb, int jn) {
  using Y = Lay<NP>;
  if constexpr (G < Y::NG) {
    constexpr int S = G & 3;
    double d[Y::NF][4];
    to_f64<NP>(d, cq[S]);
#pragma unroll
    for (int m = 0; m < Y::NF; ++m)
      if (jb + 4 * G + r0 + 2 * m >= live) d[m][0] = d[m][1] = d[m][2] = d[m][3] = 0.0;
    const int jr = G + 4 < Y::NG ? jb + 4 * (G + 4) : jn + 4 * (G + 4 - Y::NG);
#pragma unroll
    for (int m = 0; m < Y::NF; ++m) cq[S][m] = ldq<AL>(cp + min(jr + r0 + 2 * m, live - 1) * NP);
    pin();
    adj_steps<NP, 4 * G, 4 * G + 3>(h, w, d, gv);
    pin();
    adj_block<NP, AL, G + 1>(h, w, cq, gv, cp, r0, live, jb, jn);
  }
}

template <int NP, bool AL>
__device__ __forceinline__ void adj_chunk(const SumItem& it, int L, int r) {
  using Y = Lay<NP>;
  constexpr int NS = NP / 16;
  const int l = r & 15;
  const int r0 = Y::row(l, 0);
  const int live = it.live;
  const float* cp = it.Ab + Y::col(l);  // + row * NP
  double h[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) h[i] = (i == NP - 1 - r) ? 1.0 : 0.0;  // unit at position -1-r
  double w = 0.0;
  float4 cq[4][Y::NF];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int m = 0; m < Y::NF; ++m) cq[g][m] = opq(ldq<AL>(cp + min(4 * g + r0 + 2 * m, live - 1) * NP));
  float gs[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) gs[s] = opq(it.xb[min(16 * s + l, live - 1)]);
  for (int jb = 0; jb < L; jb += NP) {
    const int jn = jb + NP < L ? jb + NP : jb;
    double gv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      gv[s] = jb + 16 * s + l < live ? (double)gs[s] : 0.0;
      gs[s] = it.xb[min(jn + 16 * s + l, live - 1)];
    }
    adj_block<NP, AL, 0>(h, w, cq, gv, cp, r0, live, jb, jn);
  }
  // Psi[r][k] = h_r(L-1-k), in slot NP-1-k
#pragma unroll
  for (int k = 0; k < NP; ++k) it.out[r * NP + k] = h[NP - 1 - k];
  it.out[NP * NP + r] = w;
}

template <int NP>
__global__ __launch_bounds__(THREADS) void row_adj_summary(ChunkArgs<float> a) {
  const int64_t item = ((int64_t)blockIdx.x * THREADS + threadIdx.x) / NP;
  const int r = (int)(threadIdx.x & (NP - 1));
  SumItem it;
  if (!summary_item<NP, false>(a, item, it)) return;  // a chunk's lanes leave together
  if ((reinterpret_cast<uintptr_t>(it.Ab) & 15) == 0)
    adj_chunk<NP, true>(it, a.L, r);
  else
    adj_chunk<NP, false>(it, a.L, r);
}

}  // namespace rsum

// orders 2, 4 and 8: one quad per chunk (tlpc_quad_rerun.hip)
int launch_quad_summary(ChunkArgs<float> a, int np, bool fwd, hipStream_t st);

bool row_summary_ok(const ChunkArgs<float>& a, int np) {
  return (np == 2 || np == 4 || np == 8 || np == 16 || np == 32) && a.n == np && !a.neg &&
         (a.rg != nullptr || a.a_tstride == a.n) && a.L % (np < 16 ? 16 : np) == 0 && a.L >= np;
}

int launch_row_summary(ChunkArgs<float> a, int np, bool fwd, hipStream_t st) {
  if (np <= 8) return launch_quad_summary(a, np, fwd, st);
  const int64_t items = a.rg ? a.G : a.B * (a.P - 1);
  const int64_t blocks = (items * np + rsum::THREADS - 1) / rsum::THREADS;
  if (blocks == 0) return 0;
  const dim3 g((unsigned)blocks), b(rsum::THREADS);
  switch (np) {
    case 16:
      if (fwd) hipLaunchKernelGGL(rsum::row_fwd_summary<16>, g, b, 0, st, a);
      else hipLaunchKernelGGL(rsum::row_adj_summary<16>, g, b, 0, st, a);
      break;
    case 32:
      if (fwd) hipLaunchKernelGGL(rsum::row_fwd_summary<32>, g, b, 0, st, a);
      else hipLaunchKernelGGL(rsum::row_adj_summary<32>, g, b, 0, st, a);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // namespace tlpc
