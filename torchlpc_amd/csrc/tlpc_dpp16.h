// Order-16 building blocks on CDNA4 data-parallel primitives (DPP), shared by
// the per-signal kernel and the order-16 chunked-scan kernels.
//
// Matrix rows: a 16-lane DPP row advances 16 independent float64 recurrences
// (one per lane) through a chunk. Lane i of the row holds coefficient A[t, i+1]
// of the current step and every lane reads it through a row_newbcast operand of
// v_fmac_f64: one instruction per lag, no LDS traffic.
//
// Quads: four lanes share one serial recurrence; lane ql owns lags 4ql+1..4ql+4
// (a float4 of the A row per step, so a quad loads a whole 64-byte row) and a
// 4-slot window of the state; partial sums meet through quad_perm DPP.
#pragma once

#include "tlpc_common.h"

namespace tlpc {
namespace dpp16 {

constexpr int NP = 16;
constexpr int PF = 16;  // rows in flight per quad lane (one 16-step block)

// ---------------------------------------------------------------------------
// Adjoint matrix step (slot J = t mod 16 of the pending-sum window p,
// p[u mod 16] = pending sum of position u), the transposed recurrence:
//   lambda = p[J];  p[(J - i) & 15] -= A[t,i] * lambda   (i = 1..16)
// where i = 16 lands on slot J itself, which now stands for position t-16 and
// starts from zero (`fresh`). Returns lambda. The 16 products are independent of
// each other: the next step (slot J-1) waits only for the first. One asm
// statement, opened by s_nop 1 for the conversion that produced `a`.
template <int J>
__device__ __forceinline__ double adj_row_step(double (&p)[NP], double a) {
  const double lam = p[J];
  double fresh = 0.0;
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, -%16, %17 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%16, %17 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, -%16, %17 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, -%16, %17 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %4, -%16, %17 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %5, -%16, %17 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %6, -%16, %17 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %7, -%16, %17 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %8, -%16, %17 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %9, -%16, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %10, -%16, %17 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %11, -%16, %17 row_newbcast:11 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %12, -%16, %17 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %13, -%16, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %14, -%16, %17 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %15, -%16, %17 row_newbcast:15 row_mask:0xf bank_mask:0xf"
      : "+v"(p[(J - 1) & 15]), "+v"(p[(J - 2) & 15]), "+v"(p[(J - 3) & 15]), "+v"(p[(J - 4) & 15]),
        "+v"(p[(J - 5) & 15]), "+v"(p[(J - 6) & 15]), "+v"(p[(J - 7) & 15]), "+v"(p[(J - 8) & 15]),
        "+v"(p[(J - 9) & 15]), "+v"(p[(J - 10) & 15]), "+v"(p[(J - 11) & 15]), "+v"(p[(J - 12) & 15]),
        "+v"(p[(J - 13) & 15]), "+v"(p[(J - 14) & 15]), "+v"(p[(J - 15) & 15]), "+v"(fresh)
      : "v"(a), "v"(lam));
  p[J] = fresh;
  return lam;
}

// The same step with the coefficients of four consecutive steps held as float4
// columns: for step J, lane 4*(J & 3) + (k >> 2) of the row holds A[t, k] in
// component k & 3 of d (its float4 converted to float64). A 16-lane row then
// loads a 4-step group of rows (256 contiguous bytes) with one float4 per lane.
template <int J>
__device__ __forceinline__ double adj_row_step4(double (&p)[NP], const double (&d)[4]) {
  constexpr int L0 = 4 * (J & 3);
  const double lam = p[J];
  double fresh = 0.0;
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, -%16, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%17, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, -%18, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, -%19, %20 row_newbcast:%21 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %4, -%16, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %5, -%17, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %6, -%18, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %7, -%19, %20 row_newbcast:%22 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %8, -%16, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %9, -%17, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %10, -%18, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %11, -%19, %20 row_newbcast:%23 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %12, -%16, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %13, -%17, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %14, -%18, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %15, -%19, %20 row_newbcast:%24 row_mask:0xf bank_mask:0xf"
      : "+v"(p[(J - 1) & 15]), "+v"(p[(J - 2) & 15]), "+v"(p[(J - 3) & 15]), "+v"(p[(J - 4) & 15]),
        "+v"(p[(J - 5) & 15]), "+v"(p[(J - 6) & 15]), "+v"(p[(J - 7) & 15]), "+v"(p[(J - 8) & 15]),
        "+v"(p[(J - 9) & 15]), "+v"(p[(J - 10) & 15]), "+v"(p[(J - 11) & 15]), "+v"(p[(J - 12) & 15]),
        "+v"(p[(J - 13) & 15]), "+v"(p[(J - 14) & 15]), "+v"(p[(J - 15) & 15]), "+v"(fresh)
      : "v"(d[0]), "v"(d[1]), "v"(d[2]), "v"(d[3]), "v"(lam), "i"(L0), "i"(L0 + 1), "i"(L0 + 2),
        "i"(L0 + 3));
  p[J] = fresh;
  return lam;
}

// acc += v(lane K of the row) * w: a row broadcast of a per-lane value. The
// s_nop 1 covers a v produced by the VALU right before (a DPP source needs two
// wait states).
template <int K>
__device__ __forceinline__ void fma_bcast(double& acc, double v, double w) {
  asm("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc)
      : "v"(v), "v"(w), "i"(K));
}

// Carry link on lanes 0-15 (lane r owns s_r): returns z_r + sum_k M[r][k] s_k in
// float64, with s_k reaching every lane through a row_newbcast operand. The link
// is the serial step of the carry chain, so its latency is what counts: four
// accumulators of four products each, then two levels of adds. s comes from the
// VALU at the end of the previous link, hence the s_nop 1.
__device__ __forceinline__ double chain_link(double s, const double (&m)[NP], double z) {
  double acc0 = z, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, %4, %5 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, %4, %6 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %4, %7 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %4, %8 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %4, %9 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, %4, %10 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %4, %11 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %4, %12 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %4, %13 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, %4, %14 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %4, %15 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %4, %16 row_newbcast:11 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %4, %17 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, %4, %18 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %4, %19 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %4, %20 row_newbcast:15 row_mask:0xf bank_mask:0xf"
      : "+v"(acc0), "+v"(acc1), "+v"(acc2), "+v"(acc3)
      : "v"(s), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]),
        "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "v"(m[12]), "v"(m[13]),
        "v"(m[14]), "v"(m[15]));
  return (acc0 + acc1) + (acc2 + acc3);
}

// Carry link of an 8-state chain on lanes 0-7 of each 16-lane row (lanes 8-15
// repeat lanes 0-7): returns z_r + sum_k M[r][k] s_k, with s_k reaching every
// lane through a row_newbcast operand (lane k); four accumulators of two
// products each.
__device__ __forceinline__ double chain_link8(double s, const double (&m)[8], double z) {
  double acc0 = z, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, %4, %5 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, %4, %6 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %4, %7 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %4, %8 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, %4, %9 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, %4, %10 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %4, %11 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %4, %12 row_newbcast:7 row_mask:0xf bank_mask:0xf"
      : "+v"(acc0), "+v"(acc1), "+v"(acc2), "+v"(acc3)
      : "v"(s), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]));
  return (acc0 + acc1) + (acc2 + acc3);
}

// v of lane l ^ 16 (the other 16-lane row of the lane's 32-lane half), for a
// lane whose row parity is h = (lane >> 4) & 1, without the LDS crossbar:
// v_permlane16_swap exchanges the odd rows of its first operand with the even
// rows of its second, so with v in both, the first result holds row 0's values
// in row 1 (and row 2's in row 3) and the second holds row 1's in row 0.
__device__ __forceinline__ double xor16(double v, int h) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
  const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  const unsigned ol = h ? rl[0] : rl[1], oh = h ? rh[0] : rh[1];
  return __builtin_bit_cast(double, ((unsigned long long)oh << 32) | ol);
}

// One link of a 32-state chain inside a 32-lane half-wave, without LDS: lane
// (row half h = (lane >> 4) & 1, position l = lane & 15) holds state element
// e = 16h + l in `t` and the columns 16h..16h+15 of matrix rows l (`ma`) and
// l + 16 (`mb`). Each lane forms the partial sums of rows l and l+16 over its
// half of the columns with v_fmac_f64 reading t through row_newbcast (lane j of
// its 16-lane row), then swaps one partial with the lane 16 away. Returns
// sum_m M[e][m] t[m]. Every lane of the wave must take part.
__device__ __forceinline__ double dpp_matvec32(double t, const double (&ma)[16], const double (&mb)[16], int h) {
  double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
#define TLPC_DPP2(J0, J1, X0, X1)                                                                        \
  "v_fmac_f64_dpp %0, %2, %" #X0 " row_newbcast:" #J0 " row_mask:0xf bank_mask:0xf\n\t"                 \
  "v_fmac_f64_dpp %1, %2, %" #X1 " row_newbcast:" #J1 " row_mask:0xf bank_mask:0xf\n\t"
  asm volatile("s_nop 1\n\t" TLPC_DPP2(0, 1, 3, 4) TLPC_DPP2(2, 3, 5, 6) TLPC_DPP2(4, 5, 7, 8)
                   TLPC_DPP2(6, 7, 9, 10) TLPC_DPP2(8, 9, 11, 12) TLPC_DPP2(10, 11, 13, 14)
                       TLPC_DPP2(12, 13, 15, 16) TLPC_DPP2(14, 15, 17, 18)
               : "+v"(a0), "+v"(a1)
               : "v"(t), "v"(ma[0]), "v"(ma[1]), "v"(ma[2]), "v"(ma[3]), "v"(ma[4]), "v"(ma[5]), "v"(ma[6]),
                 "v"(ma[7]), "v"(ma[8]), "v"(ma[9]), "v"(ma[10]), "v"(ma[11]), "v"(ma[12]), "v"(ma[13]),
                 "v"(ma[14]), "v"(ma[15]));
  asm volatile("s_nop 1\n\t" TLPC_DPP2(0, 1, 3, 4) TLPC_DPP2(2, 3, 5, 6) TLPC_DPP2(4, 5, 7, 8)
                   TLPC_DPP2(6, 7, 9, 10) TLPC_DPP2(8, 9, 11, 12) TLPC_DPP2(10, 11, 13, 14)
                       TLPC_DPP2(12, 13, 15, 16) TLPC_DPP2(14, 15, 17, 18)
               : "+v"(b0), "+v"(b1)
               : "v"(t), "v"(mb[0]), "v"(mb[1]), "v"(mb[2]), "v"(mb[3]), "v"(mb[4]), "v"(mb[5]), "v"(mb[6]),
                 "v"(mb[7]), "v"(mb[8]), "v"(mb[9]), "v"(mb[10]), "v"(mb[11]), "v"(mb[12]), "v"(mb[13]),
                 "v"(mb[14]), "v"(mb[15]));
#undef TLPC_DPP2
  const double pa = a0 + a1, pb = b0 + b1;
  // element l (h = 0) takes the upper half's partial of row l; element 16 + l
  // (h = 1) the lower half's partial of row l + 16
  const double other = xor16(h ? pa : pb, h);
  return (h ? pb : pa) + other;
}

// ---------------------------------------------------------------------------
// Quad helpers.
template <int CTRL>
__device__ __forceinline__ float qperm(float v) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double qperm(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffll), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
constexpr int QP_SWAP1 = 0xB1;  // quad_perm:[1,0,3,2]
constexpr int QP_SWAP2 = 0x4E;  // quad_perm:[2,3,0,1]
constexpr int QP_UP = 0x90;     // quad_perm:[0,0,1,2]  lane ql reads lane ql-1
constexpr int QP_DOWN = 0xF9;   // quad_perm:[1,2,3,3]  lane ql reads lane ql+1
constexpr int QP_LANE0 = 0x00;  // quad_perm:[0,0,0,0]

template <typename V>
__device__ __forceinline__ V quad_sum(V p) {
  p += qperm<QP_SWAP1>(p);
  p += qperm<QP_SWAP2>(p);
  return p;
}

// Forward quad step. Lane ql's window R holds y_u for u in [t-4-4ql, t-1-4ql] at
// slot u mod 4; c = A[t, 4ql .. 4ql+3]. Returns y_t (in every lane of the quad).
template <int J, typename V>
__device__ __forceinline__ V fwd_quad_step(V (&R)[4], const float4 c, float x, int ql) {
  const float cc[4] = {c.x, c.y, c.z, c.w};
  V p = (V)cc[(J - 1) & 3] * R[0];
  p += (V)cc[(J - 2) & 3] * R[1];
  p += (V)cc[(J - 3) & 3] * R[2];
  p += (V)cc[(J - 4) & 3] * R[3];
  const V y = (V)x - quad_sum(p);
  const V up = qperm<QP_UP>(R[J]);  // lane ql-1 holds y_{t-4ql} in slot J
  R[J] = (ql == 0) ? y : up;
  return y;
}

// Adjoint quad step at time t (slot J = t mod 4), descending t. State: `now` in
// lane 0 is the pending sum of position t; lane ql's window P holds the pending
// sums of positions [t-4-4ql, t-1-4ql] at slot u mod 4. Returns lambda_t and
// leaves the state for step t-1. A step that is not `live` (past the end of the
// signal) has lambda_t = 0 and leaves the pending sums as they are.
template <int J, typename V>
__device__ __forceinline__ V adj_quad_step(V (&P)[4], V& now, const float4 c, float g, int ql,
                                           bool live) {
  const V lam0 = qperm<QP_LANE0>((V)g + now);
  const V lam = live ? lam0 : (V)0;
  const float cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
  for (int m = 0; m < 4; ++m) P[(J - 1 - m) & 3] -= (V)cc[m] * lam;
  constexpr int S = (J - 1) & 3;  // slot that leaves at the top and enters at the bottom
  const V down = qperm<QP_DOWN>(P[S]);
  now = P[S];  // meaningful in lane 0: pending sum of position t-1
  P[S] = (ql == 3) ? (V)0 : down;
  return lam;
}

}  // namespace dpp16
}  // namespace tlpc
