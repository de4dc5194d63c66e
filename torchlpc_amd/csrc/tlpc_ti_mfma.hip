// Time-invariant all-pole filters of order 16..32 (float32 data), one workgroup
// per signal, with the block products on MFMA.
//
// A signal is cut into blocks of LB = 64 samples. With h the impulse response of
// 1 / (1 + a_1 z^-1 + ... + a_N z^-N) and the state s_p of block p (the N outputs
// before it, s_p[k] = y_{64p-1-k}), the block's outputs are
//
//   y_p = H x_p + Psi s_p,        H[i][k] = h_{i-k} (lower-triangular Toeplitz),
//                                 Psi[i][k] = response at i to the unit state e_k,
//
// and the states follow s_{p+1} = Phi s_p + c_p with Phi[r][k] = Psi[63-r][k] and
// c_p[r] = (H x_p)[63-r], the zero-state tail. Per segment of 64 blocks:
//
//   P1  tails c_p = Htail x_p for all blocks          (f64 MFMA 16x16x4)
//   P2  superblock carries C_k: 8 links of Phi from 0 (f64 DPP, one superblock per half-wave)
//   P3  the chain s_{8(k+1)} = Phi^8 s_{8k} + C_k     (f64 DPP, one wave, 8 links)
//   P4  block states s_p from the superblock states   (f64 DPP, one superblock per half-wave)
//   P5  outputs Y = H X (f32 MFMA) + Psi S (f64 MFMA: the basis combination)
//
// Precision. Poles near the unit circle (radius 1 - 1e-6 at order 32 with
// reflection coefficients up to 0.8) amplify any error that enters the state.
// Measured on such signals: float32 state transitions lose 1e-4, float32 tails
// (one 64-term float32 dot product per state component, fed into the chain)
// 6e-5, and a float32 basis combination 2e-5 (|Psi| reaches several hundred,
// so Psi s cancels); H x in float32 costs 5e-8. Hence float64 for everything
// that touches the state, float32 MFMA only for the zero-state product H x.
// h, Psi, Phi and Phi^8 are built per signal in float64: Psi's first 32 rows by
// the recurrence (one lane per column), the rest by a doubling product, h from
// its first column and Phi^8 by three squarings (float64 MFMA).
//
// The adjoint of a time-invariant filter is the same filter run backwards in
// time: lambda'_t' = g'_t' - sum_i a_i lambda'_{t'-i} with g'_t' = g_{T-1-t'} and
// zero initial state, so the backward is this kernel on the reversed gradient.
// It adds dx = lambda, dzi_k = -sum_{i>k} a_i lambda_{i-1-k}, and the fused
// reduction da_i = -sum_t lambda_t y_{t-i}, formed on MFMA as
//
//   M[r][m] = sum_q lambda'_{32q+r} y'_{32q+m}   (y'_u = y_{T-1-u}, zi past the end)
//   da_i    = -sum_r M[r][r+i],
//
// with the float32 MFMA partial sums of each 2048-sample tile added into float64.
// Reference: the time-invariant gradient of torchlpc/README.md:116-128 and the
// filter of torchlpc/csrc/scan_cpu.cpp:80-122 (per-sample, serial).
#include <hip/hip_runtime.h>

#include "../../include/torchlpc_amd.h"
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"

namespace tlpc {
namespace {

using dpp16::dpp_matvec32;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int NP = 32;             // state size (orders 16..32, zero-padded coefficients)
constexpr int LB = 64;             // samples per block
constexpr int SBK = 8;             // blocks per superblock
constexpr int SEGB = 64;           // blocks per segment
constexpr int SEG = SEGB * LB;     // samples per segment
constexpr int QS = SEGB / SBK;     // superblocks per segment
constexpr int CT = 32;             // blocks per column tile
constexpr int CTS = CT * LB;       // samples per column tile
constexpr int NT = 256;            // threads per workgroup (4 waves)
constexpr int YWN = CTS + LB;      // y' window of a column tile (samples)
static_assert(QS == 2 * (NT / 64), "P2 / P4 give each 32-lane half one superblock");

// LDS pitches (in elements). The float32 MFMA operand reads have 16 consecutive
// lanes along one axis and 2 (of 4) values along the other within each 32-lane
// half: a pitch of 2 mod 32 (one axis steps by 1) or 16 mod 32 (rows of 32
// samples) keeps the 32 addresses of a half on distinct banks.
constexpr int XP = 66;   // X staging [block][sample], float
constexpr int WP = 48;   // rows of 32 samples: lambda' and the y' window, float
constexpr int DP = 33;   // rows of 32 doubles: tails / states, Psi, Phi, Phi^8

// Byte offsets of the LDS regions. The squarings of Phi use XS and YW as
// temporaries before the segments start; the da matrix reuses CS at the end.
constexpr int CS_OFF = 0;
constexpr int CS_BYTES = SEGB * DP * 8;                 // 16896 tails, then block states
constexpr int XS_OFF = CS_OFF + CS_BYTES;
constexpr int XS_BYTES = 64 * WP * 4;                   // 12288 (X staging 8448, lambda' 12288)
constexpr int YW_OFF = XS_OFF + XS_BYTES;
constexpr int YW_BYTES = ((YWN + 31) / 32) * WP * 4;    // 12672 (also P2's vectors, 2048)
constexpr int PS_OFF = YW_OFF + YW_BYTES;
constexpr int PS_BYTES = LB * DP * 8;                   // 16896
constexpr int HZ_OFF = PS_OFF + PS_BYTES;               // h, float, zero below index 64
constexpr int HD_OFF = HZ_OFF + 128 * 4;                // h, double, zero below index 64
constexpr int PH_OFF = HD_OFF + 128 * 8;
constexpr int PH_BYTES = NP * DP * 8;                   // 8448
constexpr int P8_OFF = PH_OFF + PH_BYTES;               // Phi^8
constexpr int SB_OFF = P8_OFF + PH_BYTES;
constexpr int SB_BYTES = (QS + 1) * NP * 8;             // 2304
constexpr int CU_OFF = SB_OFF + SB_BYTES;               // segment-start state, double[32]
constexpr int AC_OFF = CU_OFF + NP * 8;                 // coefficients a_0..a_63, double
constexpr int L0_OFF = AC_OFF + 64 * 8;                 // lambda_0..lambda_31, float
constexpr int LDS_BYTES = L0_OFF + NP * 4;
static_assert(NP * 34 * 8 <= XS_BYTES && NP * 34 * 8 <= YW_BYTES, "Phi powers must fit");
static_assert(CT * XP * 4 <= XS_BYTES, "X staging must fit");
static_assert(QS * NP * 8 <= YW_BYTES, "P2 vectors must fit");
static_assert(NP * 64 * 8 <= CS_BYTES, "da matrix must fit the tail region");
static_assert(2 * LDS_BYTES <= 160 * 1024, "two workgroups per compute unit");

__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// D[i][j] += sum_k A[i][k] B[k][j] in float64: lane l gives A[l%16][l/16] and
// B[l/16][l%16] and holds D[l/16 + 4r][l%16] in element r (the float32 form
// holds D[4(l/16) + r][l%16]).
__device__ __forceinline__ f64x4 mfma16d(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Columns 16h..16h+15 of rows l and l + 16 of a 32x32 double matrix (pitch DP).
__device__ __forceinline__ void load_rows32(double (&ma)[16], double (&mb)[16], const double* M, int l, int h, int pitch) {
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    ma[j] = M[l * pitch + 16 * h + j];
    mb[j] = M[(l + 16) * pitch + 16 * h + j];
  }
}

// 32x32 double product C = A B on float64 MFMA, wave w computing the 16x16 tile
// (w / 2, w % 2); row m of B is read at Bm + m * bstep (bstep may be negative,
// to walk rows backwards). Pitches in doubles.
__device__ __forceinline__ void matmul32(double* C, int cp, const double* A, int ap, const double* Bm, int bstep,
                                         int wave, int l16, int lq) {
  const int ti = wave >> 1, tj = wave & 1;
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < NP / 4; ++s) {
    const int k = 4 * s + lq;
    acc = mfma16d(A[(16 * ti + l16) * ap + k], Bm[k * bstep + 16 * tj + l16], acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) C[(16 * ti + lq + 4 * r) * cp + 16 * tj + l16] = acc[r];
}
constexpr int MP = 34;  // pitch of the Phi-power temporaries

// The MFMA A operands of the output products are fixed per wave: wave w owns
// row tiles (RB0, RB1) = (0, 3) or (1, 2) (both 20 k-steps of H) of column tile
// w / 2. They are read from LDS once and kept in registers.
template <int RB0, int RB1>
__device__ __forceinline__ void load_out_operands(float (&hop)[20], double (&pop)[16], const float* hz,
                                                  const double* psi, int l16, int lq, int prow16) {
  static_assert(4 * (RB0 + 1) + 4 * (RB1 + 1) == 20, "20 k-steps of H per wave");
#pragma unroll
  for (int s = 0; s < 4 * (RB0 + 1); ++s) hop[s] = hz[64 + 16 * RB0 + l16 - (4 * s + lq)];
#pragma unroll
  for (int s = 0; s < 4 * (RB1 + 1); ++s) hop[4 * (RB0 + 1) + s] = hz[64 + 16 * RB1 + l16 - (4 * s + lq)];
#pragma unroll
  for (int s = 0; s < NP / 4; ++s) {
    pop[s] = psi[(16 * RB0 + prow16) * DP + 4 * s + lq];
    pop[8 + s] = psi[(16 * RB1 + prow16) * DP + 4 * s + lq];
  }
}

// The two output tiles of a wave: H X (float32 MFMA) + Psi S (float64 MFMA).
template <int RB0, int RB1>
__device__ __forceinline__ void out_tiles(float (&out)[2][4], const float (&hop)[20], const double (&pop)[16],
                                          const float* xs, const double* st, int j, int lq) {
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    constexpr int n0 = 4 * (RB0 + 1);
    const int ns = q == 0 ? n0 : 4 * (RB1 + 1);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 16; ++s)
      if (s < ns) acc = mfma16(hop[(q == 0 ? 0 : n0) + s], xs[j * XP + 4 * s + lq], acc);
    f64x4 accd = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NP / 4; ++s) accd = mfma16d(pop[8 * q + s], st[j * DP + 4 * s + lq], accd);
#pragma unroll
    for (int r = 0; r < 4; ++r) out[q][r] = (float)((double)acc[r] + accd[r]);
  }
}

template <bool BWD>
__global__ __launch_bounds__(NT, 2) void ti_mfma(ChunkArgs<float> a) {
  __shared__ __align__(16) unsigned char lds[LDS_BYTES];
  double* cs = reinterpret_cast<double*>(lds + CS_OFF);
  float* xs = reinterpret_cast<float*>(lds + XS_OFF);
  float* yw = reinterpret_cast<float*>(lds + YW_OFF);
  double* psi = reinterpret_cast<double*>(lds + PS_OFF);
  float* hz = reinterpret_cast<float*>(lds + HZ_OFF);
  double* hd = reinterpret_cast<double*>(lds + HD_OFF);
  double* phi = reinterpret_cast<double*>(lds + PH_OFF);
  double* ph8 = reinterpret_cast<double*>(lds + P8_OFF);
  double* ssb = reinterpret_cast<double*>(lds + SB_OFF);
  double* cur = reinterpret_cast<double*>(lds + CU_OFF);
  double* acoef = reinterpret_cast<double*>(lds + AC_OFF);
  float* lam0 = reinterpret_cast<float*>(lds + L0_OFF);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int64_t b = blockIdx.x;
  const int64_t T = a.T;
  const int n = a.n;
  const float* src = a.x;  // forward: x; backward: the gradient g
  const int64_t base = b * T;

  // ---- setup: coefficients, Psi, h, Phi, Phi^8 (float64) ----
  if (tid < 64) acoef[tid] = (tid >= 1 && tid <= n) ? (double)a.A[b * a.a_bstride + tid - 1] : 0.0;
  if (tid < NP) {
    cur[tid] = (!BWD && a.zi && tid < n) ? (double)a.zi[b * n + tid] : 0.0;
    lam0[tid] = 0.0f;
  }
  __syncthreads();
  if (wave == 0 && lane < NP) {
    // column k of Psi, rows 0..31: psi_t = -(a_{t+1+k} + sum_{i=1..t} a_i psi_{t-i})
    // in registers; the newest term (i = 1) is added last, so a step waits on
    // its predecessor for one multiply-add only
    const int k = lane;
    double ar[NP], hist[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) ar[i] = acoef[i + 1];
#pragma unroll
    for (int t = 0; t < NP; ++t) {
      double s0 = acoef[t + 1 + k], s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
      for (int i = t; i >= 2; --i) {
        const double v = ar[i - 1] * hist[t - i];
        if ((i & 3) == 0) s0 += v;
        else if ((i & 3) == 1) s1 += v;
        else if ((i & 3) == 2) s2 += v;
        else s3 += v;
      }
      double tot = (s0 + s1) + (s2 + s3);
      if (t >= 1) tot = fma(ar[0], hist[t - 1], tot);
      hist[t] = -tot;
      psi[t * DP + k] = hist[t];
    }
  }
  __syncthreads();
  // rows 32..63: Psi[32+t][k] = sum_m Psi[t][m] Psi[31-m][k] (the state after
  // 32 steps is s[m] = y_{31-m})
  matmul32(psi + NP * DP, DP, psi, DP, psi + (NP - 1) * DP, -DP, wave, l16, lq);
  __syncthreads();  // h reads rows 32..62
  if (tid < 128) {
    // h_0 = 1, h_t = Psi[t-1][0]: the impulse leaves the unit state e_0 behind it
    const double h = tid < 64 ? 0.0 : (tid == 64 ? 1.0 : psi[(tid - 65) * DP]);
    hd[tid] = h;
    hz[tid] = (float)h;
  }
  __syncthreads();
  double* phiA = reinterpret_cast<double*>(lds + XS_OFF);  // [32][MP] Phi, then Phi^4
  double* phiB = reinterpret_cast<double*>(lds + YW_OFF);  // Phi^2, then Phi^8
  for (int e = tid; e < NP * NP; e += NT) {
    // Phi[r][k] = Psi[63-r][k]
    const int r = e >> 5, k = e & 31;
    const double v = psi[(LB - 1 - r) * DP + k];
    phi[r * DP + k] = v;
    phiA[r * MP + k] = v;
  }
  __syncthreads();
  matmul32(phiB, MP, phiA, MP, phiA, MP, wave, l16, lq);  // Phi^2
  __syncthreads();
  matmul32(phiA, MP, phiB, MP, phiB, MP, wave, l16, lq);  // Phi^4
  __syncthreads();
  matmul32(phiB, MP, phiA, MP, phiA, MP, wave, l16, lq);  // Phi^8
  __syncthreads();
  for (int e = tid; e < NP * NP; e += NT) ph8[(e >> 5) * DP + (e & 31)] = phiB[(e >> 5) * MP + (e & 31)];
  // chains in 32-lane halves: lane position cl, row half ch, state element ce
  const int cl = lane & 15, ch = (lane >> 4) & 1, ce = 16 * ch + cl;
  // rows of the float64 basis product, permuted so that its accumulator holds
  // the elements the float32 accumulator of H x holds
  const int prow16 = 4 * (l16 & 3) + (l16 >> 2);
  __syncthreads();

  // per-wave MFMA A operands: Htail for the tails, H and Psi rows for the outputs
  // (the backward, which also holds the da accumulators and the y' prefetch,
  // reads Htail from LDS)
  constexpr bool kHtailRegs = !BWD;
  double htop[kHtailRegs ? LB / 4 : 1];
  if constexpr (kHtailRegs) {
#pragma unroll
    for (int s = 0; s < LB / 4; ++s) htop[s] = hd[127 - (16 * (wave >> 1) + l16) - (4 * s + lq)];
  }
  float hop[20];
  double pop[16];
  if (wave & 1)
    load_out_operands<1, 2>(hop, pop, hz, psi, l16, lq, prow16);
  else
    load_out_operands<0, 3>(hop, pop, hz, psi, l16, lq, prow16);

  // da accumulators (backward): M tiles (row tile mrt, column tiles mct0, mct0+1)
  double macc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  const int mrt = wave & 1, mct0 = 2 * (wave >> 1);

  const int64_t P = (T + LB - 1) / LB;
  const int64_t nsegs = (P + SEGB - 1) / SEGB;

  // x tiles (g tiles backward) are needed in the order P1 tiles of a segment,
  // then its P5 tiles, then the next segment's; each is loaded into registers
  // one step ahead (during the previous tile's products) and written to the
  // staging region when its turn comes.
  float pre[CTS / NT];
  auto load_tile = [&](int64_t t0) {
#pragma unroll
    for (int e = 0; e < CTS / NT; ++e) {
      const int64_t t = t0 + e * NT + tid;
      float v = 0.0f;
      if (t < T) v = BWD ? src[base + T - 1 - t] : src[base + t];
      pre[e] = v;
    }
  };
  auto stage_tile = [&]() {
#pragma unroll
    for (int e = 0; e < CTS / NT; ++e) {
      const int s = e * NT + tid;
      xs[(s >> 6) * XP + (s & 63)] = pre[e];
    }
  };
  constexpr int YPT = (YWN + NT - 1) / NT;
  float prey[YPT];
  auto load_ywin = [&](int64_t t0) {
    // y'_u = y_{T-1-u} for u < T, zi[u-T] past the end (zero past the order)
#pragma unroll
    for (int e = 0; e < YPT; ++e) {
      const int s = e * NT + tid;
      const int64_t u = t0 + s;
      float v = 0.0f;
      if (s < YWN) {
        if (u < T)
          v = a.y[base + T - 1 - u];
        else if (u - T < n && a.zi)
          v = a.zi[b * n + (u - T)];
      }
      prey[e] = v;
    }
  };
  auto stage_ywin = [&]() {
#pragma unroll
    for (int e = 0; e < YPT; ++e) {
      const int s = e * NT + tid;
      if (s < YWN) yw[(s >> 5) * WP + (s & 31)] = prey[e];
    }
  };
  const bool with_da = BWD && a.da_ti;

  load_tile(0);
  for (int64_t seg = 0; seg < nsegs; ++seg) {
    const int64_t t_seg = seg * SEG;
    const int nb = (int)((P - seg * SEGB) < SEGB ? (P - seg * SEGB) : SEGB);
    const int nct = (nb + CT - 1) / CT;
    const int nsb = (nb + SBK - 1) / SBK;

    // ---- P1: tails, c_p[r] = sum_k h_{63-r-k} x_{64p+k} ----
#pragma unroll 1
    for (int ct = 0; ct < nct; ++ct) {
      stage_tile();
      __syncthreads();
      // next: this segment's next P1 tile, else its first P5 tile
      load_tile(t_seg + (ct + 1 < nct ? (int64_t)(ct + 1) * CTS : 0));
      {
        const int rr0 = 16 * (wave >> 1), jj0 = 16 * (wave & 1);
        f64x4 acc = {0.0, 0.0, 0.0, 0.0};
        const int i = rr0 + l16;
        const int j = jj0 + l16;
#pragma unroll
        for (int s = 0; s < LB / 4; ++s) {
          double hv;
          if constexpr (kHtailRegs)
            hv = htop[s];
          else
            hv = hd[127 - i - (4 * s + lq)];
          acc = mfma16d(hv, (double)xs[j * XP + 4 * s + lq], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) cs[(ct * CT + j) * DP + rr0 + lq + 4 * r] = acc[r];
      }
      __syncthreads();
    }
    if (with_da) load_ywin(t_seg);

    // ---- P2: superblock carries (zero start, 8 links of Phi) ----
    // wave w: superblocks 2w and 2w + 1 (one per 32-lane half), lane element ce
    {
      double ma[16], mb[16];
      load_rows32(ma, mb, phi, cl, ch, DP);
      const int k = 2 * wave + (lane >> 5);
      const bool live = k < nsb;
      double tk = 0.0;
#pragma unroll 1
      for (int j = 0; j < SBK; ++j) {
        const int blk = k * SBK + j;
        const double c = (live && blk < nb) ? cs[blk * DP + ce] : 0.0;
        const double mv = dpp_matvec32(tk, ma, mb, ch);
        tk = (j == 0 ? 0.0 : mv) + c;
      }
      if (live) ssb[k * NP + ce] = tk;
    }
    __syncthreads();

    // ---- P3: the superblock chain (wave 0; both halves carry the same chain) ----
    if (wave == 0) {
      double ma[16], mb[16];
      load_rows32(ma, mb, ph8, cl, ch, DP);
      double st = cur[ce];
#pragma unroll 1
      for (int k = 0; k < nsb; ++k) {
        const double ck = ssb[k * NP + ce];
        const double nxt = dpp_matvec32(st, ma, mb, ch) + ck;
        if (lane < 32) ssb[k * NP + ce] = st;  // the state at superblock k's start
        st = nxt;
      }
      if (lane < 32) cur[ce] = st;
    }
    __syncthreads();

    // ---- P4: block states from the superblock states ----
    {
      double ma[16], mb[16];
      load_rows32(ma, mb, phi, cl, ch, DP);
      const int k = 2 * wave + (lane >> 5);
      const bool live = k < nsb;
      double tk = live ? ssb[k * NP + ce] : 0.0;
#pragma unroll 1
      for (int j = 0; j < SBK; ++j) {
        const int blk = k * SBK + j;
        double c = 0.0;
        if (live && blk < nb) {
          c = cs[blk * DP + ce];
          cs[blk * DP + ce] = tk;  // the state at block blk's start replaces its tail
        }
        tk = dpp_matvec32(tk, ma, mb, ch) + c;
      }
    }
    __syncthreads();

    // ---- P5: outputs Y = H X + Psi S (and, backward, the da products) ----
#pragma unroll 1
    for (int ct = 0; ct < nct; ++ct) {
      const int64_t t0 = t_seg + (int64_t)ct * CTS;
      stage_tile();
      if (with_da) stage_ywin();
      __syncthreads();
      // next: this segment's next P5 tile, else the next segment's first P1 tile
      if (ct + 1 < nct) {
        load_tile(t0 + CTS);
        if (with_da) load_ywin(t0 + CTS);
      } else if (seg + 1 < nsegs) {
        load_tile(t_seg + SEG);
      }
      float out[2][4];
      // wave w: tiles (row tile rb, column tile cb): w0 (0,0)(3,0), w1 (1,0)(2,0),
      // w2 (0,1)(3,1), w3 (1,1)(2,1)
      const int cb = wave >> 1;
      if (wave & 1)
        out_tiles<1, 2>(out, hop, pop, xs, cs + ct * CT * DP, 16 * cb + l16, lq);
      else
        out_tiles<0, 3>(out, hop, pop, xs, cs + ct * CT * DP, 16 * cb + l16, lq);
      // The outputs leave through the staging region in sample order: straight
      // from the MFMA layout a store instruction would cover 16 blocks with four
      // floats each, a cache request per 16 bytes.
      if (!BWD) {
        __syncthreads();  // every wave has read the X staging
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int rb = (wave & 1) ? (q == 0 ? 1 : 2) : (q == 0 ? 0 : 3);
          const int j = 16 * cb + l16;
#pragma unroll
          for (int r = 0; r < 4; ++r) xs[j * XP + 16 * rb + 4 * lq + r] = out[q][r];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < CTS / NT; ++e) {
          const int sl = e * NT + tid;
          const int64_t t = t0 + sl;
          if (t < T) a.y[base + t] = xs[(sl >> 6) * XP + (sl & 63)];
        }
      }
      if (BWD) {
        __syncthreads();  // X staging is read; lambda' goes to the same region
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int rb = (wave & 1) ? (q == 0 ? 1 : 2) : (q == 0 ? 0 : 3);
          const int j = 16 * cb + l16;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * rb + 4 * lq + r;
            const int tl = j * LB + i;
            const int64_t t = t0 + tl;
            const float v = t < T ? out[q][r] : 0.0f;
            xs[(tl >> 5) * WP + (tl & 31)] = v;
            if (t < T && t >= T - NP) lam0[T - 1 - t] = v;
          }
        }
        __syncthreads();
        if (a.dx) {
          // dx_t = lambda_t, from the staged lambda' in sample order (descending t)
#pragma unroll
          for (int e = 0; e < CTS / NT; ++e) {
            const int tl = e * NT + tid;
            const int64_t t = t0 + tl;
            if (t < T) a.dx[base + T - 1 - t] = xs[(tl >> 5) * WP + (tl & 31)];
          }
        }
        if (with_da) {
          // M[r][m] += sum_q lambda'_{32q+r} y'_{32q+m}, q over the tile's 64 rows
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int mct = mct0 + h;
            const int r = 16 * mrt + l16;
            const int m = 16 * mct + l16;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
            for (int s = 0; s < 16; ++s) {
              const int q = 4 * s + lq;
              acc = mfma16(xs[q * WP + r], yw[(q + (m >> 5)) * WP + (m & 31)], acc);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) macc[h][e] += (double)acc[e];
          }
        }
      }
      __syncthreads();
    }
  }

  if (BWD) {
    if (a.da_ti) {
      // da_i = -sum_r M[r][r+i]; the M tiles go through the tail region
      double* mm = reinterpret_cast<double*>(lds + CS_OFF);  // [32][64]
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) mm[(16 * mrt + 4 * lq + e) * 64 + 16 * (mct0 + h) + l16] = macc[h][e];
      __syncthreads();
      if (tid < n) {
        const int i = tid + 1;
        double s = 0.0;
        for (int r = 0; r < NP; ++r) s += mm[r * 64 + r + i];
        a.da_ti[b * n + tid] = (float)(-s);
      }
    }
    if (a.dzi && tid < n) {
      // dzi_k = -sum_{i=k+1..n} a_i lambda_{i-1-k}
      double s = 0.0;
      for (int i = tid + 1; i <= n; ++i) s += acoef[i] * (double)lam0[i - 1 - tid];
      a.dzi[b * n + tid] = (float)(-s);
    }
  }
}

}  // namespace

bool ti_mfma_ok(int n, int64_t T) { return n >= 16 && n <= NP && T > 0; }

int launch_ti_mfma_fwd(ChunkArgs<float> a, hipStream_t st) {
  if (!ti_mfma_ok(a.n, a.T) || a.a_tstride != 0 || a.rg) return TLPC_EINVAL;
  hipLaunchKernelGGL((ti_mfma<false>), dim3((unsigned)a.B), dim3(NT), 0, st, a);
  return hipGetLastError() == hipSuccess ? TLPC_OK : TLPC_ELAUNCH;
}

int launch_ti_mfma_bwd(ChunkArgs<float> a, hipStream_t st) {
  if (!ti_mfma_ok(a.n, a.T) || a.a_tstride != 0 || a.rg) return TLPC_EINVAL;
  hipLaunchKernelGGL((ti_mfma<true>), dim3((unsigned)a.B), dim3(NT), 0, st, a);
  return hipGetLastError() == hipSuccess ? TLPC_OK : TLPC_ELAUNCH;
}

}  // namespace tlpc
