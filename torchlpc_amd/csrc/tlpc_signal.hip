// Per-signal fused forward filter for order 16, float32 (TLPC_STRATEGY_SIGNAL).
//
// One workgroup of 13 wavefronts owns one signal and walks it in rounds of 32
// sub-chunks of Lc samples. The waves keep fixed roles and form a three-stage
// pipeline over the rounds (iteration k):
//
//   waves 0-7   round k, transition matrices: each 16-lane DPP row owns one
//               sub-chunk; lane kk runs the homogeneous recurrence from the unit
//               state e_kk in float64, and lane i also holds coefficient A[t, i]
//               of the current step, which the row reads through a row_newbcast
//               operand of v_fmac_f64 (one instruction per lag). After Lc steps
//               lane kk holds column kk of Phi_g and stores it (as float32) in LDS.
//   waves 8-9   round k, zero-state responses z_g in float64: one 4-lane quad per
//               sub-chunk, each lane owning 4 lags (a float4 of A per step) and a
//               4-deep window of the output history; the quad all-reduces the
//               partial sums with quad_perm DPP.
//   wave 12     round k-1, carry chain s_{g+1} = Phi_g s_g + z_g in float64 on 16
//               lanes (s broadcast with row_newbcast DPP), carries to LDS.
//   waves 10-11 round k-2, re-filter: one quad per sub-chunk runs the float32
//               recurrence from its carry and writes y.
//
// Every A row is read from HBM once by the matrix waves (64-byte coalesced row
// loads); the zero-state quads re-read it in the same round (L2) and the
// re-filter quads two rounds later (Infinity Cache). Float64 for the matrices,
// zero states and carries keeps the result at the accuracy of a serial float32
// recurrence even over many sub-chunk boundaries.
//
// Reference recurrence: torchlpc/csrc/scan_cpu.cpp:112-120.
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"

namespace tlpc {
namespace sig {

using dpp16::chain_link;

constexpr int NP = 16;
constexpr int MAT_WAVES = 8;
constexpr int SUBS = MAT_WAVES * 4;             // sub-chunks per round
constexpr int QUAD_WAVES = SUBS * 4 / 64;       // waves per quad role
constexpr int ZS_WAVE0 = MAT_WAVES;             // zero-state quads
constexpr int RR_WAVE0 = ZS_WAVE0 + QUAD_WAVES; // re-filter quads
constexpr int CHAIN_WAVE = RR_WAVE0 + QUAD_WAVES;
constexpr int WAVES = CHAIN_WAVE + 1;
constexpr int THREADS = WAVES * 64;
constexpr int PF = 8;  // A rows in flight per quad lane

struct Lds {
  float phi[2][SUBS][NP][NP];  // phi[buf][g][r][k] = Phi_g[r][k]
  float z[2][SUBS][NP];
  double carry[2][SUBS][NP];   // state entering sub-chunk g: carry[.][g][r] = y_{t0-1-r}
};

// ---------------------------------------------------------------------------
// Matrix rows.
//
// One step of the homogeneous recurrence for a DPP row, in float64:
//   returns -sum_{i=1..16} A[t,i] * ring[(J - i) & 15]
// with A[t,i] (as float64) in lane i-1 of the row, read by every lane through a
// negated row_newbcast operand of v_fmac_f64. Two accumulators interleave the
// even and odd lags. The step is one asm statement (nothing is scheduled between
// its ops) and opens with s_nop 1, the two wait states a DPP source needs after
// the VALU conversion that produced it.
template <int J>
__device__ __forceinline__ double basis_step(double a, const double (&r)[NP]) {
  double acc0 = 0.0, acc1 = 0.0;
  asm("s_nop 1\n\t"
      "v_fmac_f64_dpp %0, -%2, %3 row_newbcast:15 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %4 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%2, %5 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %6 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%2, %7 row_newbcast:11 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %8 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%2, %9 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %10 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%2, %11 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %12 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%2, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %14 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%2, %15 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %16 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %0, -%2, %17 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, -%2, %18 row_newbcast:0 row_mask:0xf bank_mask:0xf"
      : "+v"(acc0), "+v"(acc1)
      : "v"(a), "v"(r[(J - 16) & 15]), "v"(r[(J - 15) & 15]), "v"(r[(J - 14) & 15]),
        "v"(r[(J - 13) & 15]), "v"(r[(J - 12) & 15]), "v"(r[(J - 11) & 15]),
        "v"(r[(J - 10) & 15]), "v"(r[(J - 9) & 15]), "v"(r[(J - 8) & 15]), "v"(r[(J - 7) & 15]),
        "v"(r[(J - 6) & 15]), "v"(r[(J - 5) & 15]), "v"(r[(J - 4) & 15]), "v"(r[(J - 3) & 15]),
        "v"(r[(J - 2) & 15]), "v"(r[(J - 1) & 15]));
  return acc0 + acc1;
}

// 16 unrolled steps of a matrix lane. cf is a ring of coefficients: step J of the
// block reads cf[J] and issues the load for step J+8 into cf[(J+8) & 15], so
// eight coefficient loads stay in flight.
template <int J>
__device__ __forceinline__ void basis_block(double (&r)[NP], float (&cf)[NP], const float* ap,
                                            int jb, int Lc, int64_t t0, int64_t T) {
  if constexpr (J < NP) {
    const int jn = jb + J + 8;
    cf[(J + 8) & 15] = (jn < Lc && t0 + jn < T) ? ap[(int64_t)jn * NP] : 0.f;
    r[J] = basis_step<J>((double)cf[J], r);
    basis_block<J + 1>(r, cf, ap, jb, Lc, t0, T);
  }
}

__device__ __noinline__ void matrix_role(const ChunkArgs<float>& a, Lds& sh, int k, int wave,
                                         int lane) {
  const int64_t T = a.T;
  const int Lc = a.L;
  const int g = wave * 4 + (lane >> 4);
  const int kk = lane & 15;
  const int64_t t0 = ((int64_t)k * SUBS + g) * Lc;
  if (t0 >= T) return;
  double r[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) r[i] = (i == NP - 1 - kk) ? 1.0 : 0.0;  // y_{-1-kk} = 1
  const float* ap = a.A + (int64_t)blockIdx.x * T * NP + t0 * NP + kk;
  float cf[NP];
#pragma unroll
  for (int j = 0; j < 8; ++j) cf[j] = (j < Lc && t0 + j < T) ? ap[j * NP] : 0.f;
  for (int jb = 0; jb < Lc; jb += NP) basis_block<0>(r, cf, ap, jb, Lc, t0, T);
  // column kk of Phi_g: Phi[rr][kk] = y_{Lc-1-rr}, slot 15-rr
  float* col = &sh.phi[k & 1][g][0][kk];
#pragma unroll
  for (int rr = 0; rr < NP; ++rr) col[rr * NP] = (float)r[NP - 1 - rr];
}

// ---------------------------------------------------------------------------
// Quads. Lane ql of a quad owns lags 4ql+1..4ql+4: a float4 of A per step and a
// 4-slot window R holding y_u for u in [t-4-4ql, t-1-4ql] at slot u mod 4.

template <int CTRL>
__device__ __forceinline__ float qperm(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double qperm(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffll), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
constexpr int QP_SWAP1 = 0xB1;   // quad_perm:[1,0,3,2]
constexpr int QP_SWAP2 = 0x4E;   // quad_perm:[2,3,0,1]
constexpr int QP_SHIFT = 0x90;   // quad_perm:[0,0,1,2]

// y_t = x_t - sum_i A[t,i] y_{t-i} for the quad, step index J = t mod 4.
template <int J, typename V>
__device__ __forceinline__ V quad_step(V (&R)[4], const float4 c, float x, int ql) {
  const float cc[4] = {c.x, c.y, c.z, c.w};
  V p = (V)cc[(J - 1 - 0) & 3] * R[0];
  p += (V)cc[(J - 1 - 1) & 3] * R[1];
  p += (V)cc[(J - 1 - 2) & 3] * R[2];
  p += (V)cc[(J - 1 - 3) & 3] * R[3];
  p += qperm<QP_SWAP1>(p);
  p += qperm<QP_SWAP2>(p);
  const V y = (V)x - p;
  const V shifted = qperm<QP_SHIFT>(R[J]);  // lane ql-1's y_{t-4ql}
  R[J] = (ql == 0) ? y : shifted;
  return y;
}

// PF unrolled steps of a quad (template recursion keeps t mod 4 static): step
// jb+JJ uses row slot JJ and refills it with the row PF steps ahead.
template <int JJ, bool STORE, typename V>
__device__ __forceinline__ void quad_block(V (&R)[4], float4 (&cq)[PF], float (&xq)[PF],
                                           const float* Ab, const float* xb, float* yb,
                                           int64_t t0, int jb, int len, int ql) {
  if constexpr (JJ < PF) {
    const int step = jb + JJ;
    if (step < len) {
      const int64_t t = t0 + step;
      const float4 c = cq[JJ];
      const float xv = xq[JJ];
      if (step + PF < len) {
        cq[JJ] = reinterpret_cast<const float4*>(Ab + (t + PF) * NP)[ql];
        xq[JJ] = xb[t + PF];
      }
      const V y = quad_step<JJ & 3>(R, c, xv, ql);
      if (STORE && ((t & 3) == ql)) yb[t] = (float)y;
    }
    quad_block<JJ + 1, STORE>(R, cq, xq, Ab, xb, yb, t0, jb, len, ql);
  }
}

// Runs `len` steps of a quad from window R, prefetching PF float4 rows ahead.
// STORE: write y (re-filter); otherwise only the final window is kept.
template <bool STORE, typename V>
__device__ __forceinline__ void quad_run(V (&R)[4], const float* Ab, const float* xb, float* yb,
                                         int64_t t0, int len, int ql) {
  float4 cq[PF];
  float xq[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    if (j < len) {
      cq[j] = reinterpret_cast<const float4*>(Ab + (t0 + j) * NP)[ql];
      xq[j] = xb[t0 + j];
    }
  }
  for (int jb = 0; jb < len; jb += PF) quad_block<0, STORE>(R, cq, xq, Ab, xb, yb, t0, jb, len, ql);
}

__device__ __noinline__ void zero_state_role(const ChunkArgs<float>& a, Lds& sh, int k, int qwave,
                                             int lane) {
  const int64_t T = a.T;
  const int Lc = a.L;
  const int g = qwave * 16 + (lane >> 2);
  const int ql = lane & 3;
  const int64_t t0 = ((int64_t)k * SUBS + g) * Lc;
  if (t0 >= T) return;
  const int len = (int)((T - t0) < Lc ? (T - t0) : Lc);
  double R[4] = {0.0, 0.0, 0.0, 0.0};
  const int64_t b = blockIdx.x;
  quad_run<false>(R, a.A + b * T * NP, a.x + b * T, nullptr, t0, len, ql);
  // z[4ql+m] = y_{t1-1-4ql-m}, slot 3-m
#pragma unroll
  for (int m = 0; m < 4; ++m) sh.z[k & 1][g][4 * ql + m] = (float)R[3 - m];
}

__device__ __noinline__ void refilter_role(const ChunkArgs<float>& a, Lds& sh, int kr, int qwave,
                                           int lane) {
  const int64_t T = a.T;
  const int Lc = a.L;
  const int g = qwave * 16 + (lane >> 2);
  const int ql = lane & 3;
  const int64_t t0 = ((int64_t)kr * SUBS + g) * Lc;
  if (t0 >= T) return;
  const int len = (int)((T - t0) < Lc ? (T - t0) : Lc);
  float R[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) R[3 - m] = (float)sh.carry[kr & 1][g][4 * ql + m];
  const int64_t b = blockIdx.x;
  quad_run<true>(R, a.A + b * T * NP, a.x + b * T, a.y + b * T, t0, len, ql);
}

// ---------------------------------------------------------------------------
// Carry chain: s' = Phi_g s + z_g on lanes 0-15 (lane r owns s_r), float64.
__device__ __noinline__ void chain_role(const ChunkArgs<float>& a, Lds& sh, int kr, int lane,
                                        double& s) {
  const int64_t T = a.T;
  const int Lc = a.L;
  const int64_t rbase = (int64_t)kr * SUBS * Lc;
  const int64_t left = (T - rbase + Lc - 1) / Lc;
  const int nsub = (int)(left < SUBS ? left : SUBS);
  const int buf = kr & 1;
  const int row = lane & 15;
  for (int g = 0; g < nsub; ++g) {
    if (lane < NP) sh.carry[buf][g][lane] = s;
    double p[NP];
    const float4* pr = reinterpret_cast<const float4*>(&sh.phi[buf][g][row][0]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float4 f = pr[v];
      p[4 * v + 0] = f.x;
      p[4 * v + 1] = f.y;
      p[4 * v + 2] = f.z;
      p[4 * v + 3] = f.w;
    }
    s = chain_link(s, p, (double)sh.z[buf][g][row]);
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(THREADS) void sig_fwd16(ChunkArgs<float> a) {
  __shared__ Lds sh;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t round_len = (int64_t)SUBS * a.L;
  const int R = (int)((a.T + round_len - 1) / round_len);

  // chain state (lanes 0-15 of the chain wave): s = y_{t0-1-lane}, float64
  double s = 0.0;
  if (wave == CHAIN_WAVE && lane < NP && a.zi) s = (double)a.zi[(int64_t)blockIdx.x * NP + lane];

  for (int k = 0; k <= R + 1; ++k) {
    if (wave < MAT_WAVES) {
      if (k < R) matrix_role(a, sh, k, wave, lane);
    } else if (wave < RR_WAVE0) {
      if (k < R) zero_state_role(a, sh, k, wave - ZS_WAVE0, lane);
    } else if (wave < CHAIN_WAVE) {
      if (k >= 2) refilter_role(a, sh, k - 2, wave - RR_WAVE0, lane);
    } else {
      if (k >= 1 && k <= R) chain_role(a, sh, k - 1, lane, s);
    }
    __syncthreads();
  }
}

}  // namespace sig

int launch_signal_fwd16(ChunkArgs<float> a, hipStream_t st) {
  hipLaunchKernelGGL(sig::sig_fwd16, dim3((unsigned)a.B), dim3(sig::THREADS), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

// Sub-chunk length: a multiple of 16 near 128, chosen so the rounds of 32
// sub-chunks cover T evenly.
int signal_chunk_len(int64_t T, int requested) {
  if (requested > 0) return (requested + 15) / 16 * 16;
  const int64_t target = 128;
  int64_t R = (T + sig::SUBS * target - 1) / (sig::SUBS * target);
  if (R < 1) R = 1;
  int64_t L = (T + sig::SUBS * R - 1) / (sig::SUBS * R);
  L = (L + 15) / 16 * 16;
  if (L < 16) L = 16;
  return (int)L;
}

}  // namespace tlpc
