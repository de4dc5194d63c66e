// Reruns (phase 3 of the chunked scan) for float32 time-varying filters of
// order 2 and 4, and chunk summaries (phase 1) of order 2, 4 and 8, in uniform
// or ragged batches,
// whatever the alignment of their streams: one quad of lanes per chunk, lane q holding sample 4k + q of every
// 4-step block k of the chunk (its x or g, its coefficient row, its entering y
// value), and writing that sample's outputs. The quad's four loads of a block
// touch consecutive addresses, one cache line (two across a line boundary),
// where one lane per chunk issues a request per stream and step: the one-lane
// reruns at these orders were bound by those requests, not by bytes. Each step
// broadcasts lane (step mod 4)'s values to the quad (quad_perm), and all four
// lanes run the chunk's serial recurrence (order 2 or 4: a few FMAs per step).
//
// Arithmetic is float64 on the float32 data, as in the other reruns. Blocks are
// prefetched RB blocks ahead at clamped times; samples past the end of a signal
// compute on clamped data, the forward stores what the last live sample stored
// and the adjoint has lambda = 0 there.
//
// Reference recurrence: torchlpc/csrc/scan_cpu.cpp:112-120; adjoint:
// torchlpc/core.py:186-229.
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"
#include "tlpc_vmem.h"

namespace tlpc {
namespace quadr {

using dpp16::qperm;
using namespace vmem;

constexpr int RB = 4;        // 4-step blocks in flight per lane
constexpr int THREADS = 64;  // 16 chunks per workgroup

// lane J of the quad's value, in every lane of it (quad_perm [J, J, J, J])
template <int J>
__device__ __forceinline__ float bq(float v) {
  return qperm<J * 0x55>(v);
}

// v[k] for k in 0..3, as two levels of selects on the bits of k (a chain of
// comparisons becomes control flow, whose joins drain the prefetch ring)
__device__ __forceinline__ float pick4(int k, float v0, float v1, float v2, float v3) {
  const bool b0 = (k & 1) != 0, b1 = (k & 2) != 0;
  const float lo = b0 ? v1 : v0;
  const float hi = b0 ? v3 : v2;
  return b1 ? hi : lo;
}

// ---------------------------------------------------------------------------
// Forward. ring[u mod NP] = y_u (chunk starts are multiples of 4, so step j of a
// block sits in slot j mod NP).
template <int NP>
struct FwdQ {
  float xq[RB];
  float aq[RB][NP];
};

template <int NP, int J>
__device__ __forceinline__ double fwd_qstep(double (&ring)[NP], float xv, const float (&row)[NP]) {
  constexpr int S = J & (NP - 1);
  double acc = (double)bq<J>(xv);
#pragma unroll
  for (int i = NP; i >= 1; --i) acc = fma(-(double)bq<J>(row[i - 1]), ring[(S - i) & (NP - 1)], acc);
  ring[S] = acc;
  return acc;
}

template <int NP, int SL>
__device__ __forceinline__ void fwd_blocks(double (&ring)[NP], FwdQ<NP>& f, float& ylast, const float* Ab,
                                           const float* xb, float* yb, int64_t T, int64_t tb, int q) {
  if constexpr (SL < RB) {
    // block at tb + 4 SL; lane q's sample tb + 4 SL + q
    const int64_t t0 = tb + 4 * SL;
    float row[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) row[i] = f.aq[SL][i];
    const float xv = f.xq[SL];
    float yout = 0.f;
    {
      const double y0 = fwd_qstep<NP, 0>(ring, xv, row);
      const double y1 = fwd_qstep<NP, 1>(ring, xv, row);
      const double y2 = fwd_qstep<NP, 2>(ring, xv, row);
      const double y3 = fwd_qstep<NP, 3>(ring, xv, row);
      const float v0 = (float)y0, v1 = (float)y1, v2 = (float)y2, v3 = (float)y3;
      yout = pick4(q, v0, v1, v2, v3);
      // the last output of the block that lies in the signal (none: keep)
      const int64_t nlive = T - t0;  // samples of the block in the signal (may be <= 0 or > 4)
      const int kl = (int)(nlive < 4 ? nlive : 4) - 1;
      const float last = pick4(kl < 0 ? 0 : kl, v0, v1, v2, v3);
      ylast = kl >= 0 ? last : ylast;
    }
    // refill the slot with the block RB ahead (times clamped into the signal)
    {
      const int64_t tn = min(t0 + 4 * RB + q, T - 1);
      f.xq[SL] = xb[tn];
#pragma unroll
      for (int i = 0; i < NP; ++i) f.aq[SL][i] = Ab[tn * NP + i];
    }
    pin();
    // a sample past the end stores the last live output at T-1 again
    const int64_t tm = t0 + q;
    const bool live = tm < T;
    yb[live ? tm : T - 1] = live ? yout : ylast;
    fwd_blocks<NP, SL + 1>(ring, f, ylast, Ab, xb, yb, T, tb, q);
  }
}

template <int NP>
__global__ __launch_bounds__(THREADS) void quad_fwd_rerun(ChunkArgs<float> a) {
  const int64_t gt = (int64_t)blockIdx.x * THREADS + threadIdx.x;
  const int64_t cid = gt >> 2;
  const int q = (int)(gt & 3);
  if (cid >= (a.rg ? a.G : a.B * a.P)) return;  // a quad's lanes leave together
  const int64_t b = chunk_signal(a, cid);
  const Span sp = sig_span(a, b);
  const int64_t t0 = (cid - sp.cbase) * a.L;
  const int64_t rem = sp.T - t0;
  const int len = rem < a.L ? (int)rem : a.L;
  const float* Ab = a.A + sp.ao;
  const float* xb = a.x + sp.xo;
  float* yb = a.y + sp.xo;
  const double* cr = a.carry + cid * NP;
  double ring[NP];
#pragma unroll
  for (int r = 0; r < NP; ++r) ring[NP - 1 - r] = cr[r];  // y_{t0-1-r}
  FwdQ<NP> f;
#pragma unroll
  for (int s = 0; s < RB; ++s) {
    const int64_t tn = min(t0 + 4 * s + q, sp.T - 1);
    f.xq[s] = opq(xb[tn]);
#pragma unroll
    for (int i = 0; i < NP; ++i) f.aq[s][i] = opq(Ab[tn * NP + i]);
  }
  settle(ring);
  float ylast = 0.f;
  for (int jb = 0; jb < len; jb += 4 * RB) fwd_blocks<NP, 0>(ring, f, ylast, Ab, xb, yb, sp.T, t0 + jb, q);
}

// ---------------------------------------------------------------------------
// Adjoint, descending. P[u mod NP] is the pending sum of position u, Y[u mod
// NP] the y value of position u for the NP positions below the current step
// (block starts are 3 mod 4, so step j of a block sits in slot (3 - j) mod NP).
template <int NP>
struct AdjQ {
  float gq[RB];
  float aq[RB][NP];
  float yq[RB];  // raw y value entering the window at the lane's sample (select at use)
};

struct AdjIo {
  const float* Ab;
  const float* gb;
  const float* yb;
  const float* zb;  // zi of the signal, or yb without zi (an always-valid address)
  float* dxb;
  float* dAb;
  float* sink;  // the chunk's consumed carry slot: destination of stores past T
  int64_t T;
  int n;
  bool haszi;
};

__device__ __forceinline__ const float* yaddr(const AdjIo& io, int64_t u) { return yw_ptr(io.yb, io.zb, io.n, io.T, u); }
__device__ __forceinline__ bool yflag(const AdjIo& io, int64_t u) { return yw_flag(io.haszi, io.n, io.T, u); }

template <int NP, bool DA, int J>
__device__ __forceinline__ float adj_qstep(double (&P)[NP], float (&Y)[NP], float (&d)[NP], float gv,
                                           const float (&row)[NP], float yv, const AdjIo& io, int64_t t) {
  constexpr int S = (3 - J) & (NP - 1);  // slot of t
  const bool live = t < io.T;
  const double lam0 = (double)bq<J>(gv) + P[S];
  const double lam = live ? lam0 : 0.0;
  P[S] = live ? 0.0 : P[S];
#pragma unroll
  for (int i = 1; i <= NP; ++i) P[(S - i) & (NP - 1)] = fma(-(double)bq<J>(row[i - 1]), lam, P[(S - i) & (NP - 1)]);
  const float lf = (float)lam;
  if constexpr (DA) {
#pragma unroll
    for (int i = 1; i <= NP; ++i) d[i - 1] = -lf * Y[(S - i) & (NP - 1)];
    // the window moves down: slot (t-1) mod NP takes y_{t-1-NP}
    const float yin = bq<J>(yv);
    Y[(S - 1) & (NP - 1)] = yflag(io, t - 1 - NP) ? yin : 0.f;
  }
  return lf;
}

template <int NP, bool DA, bool DX, int SL>
__device__ __forceinline__ void adj_blocks(double (&P)[NP], float (&Y)[NP], AdjQ<NP>& r, const AdjIo& io,
                                           int64_t tb, int q) {
  if constexpr (SL < RB) {
    // block with samples t1 - j (j = 0..3), t1 = tb - 4 SL; lane q's sample t1 - q
    const int64_t t1 = tb - 4 * SL;
    float row[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) row[i] = r.aq[SL][i];
    const float gv = r.gq[SL], yv = r.yq[SL];
    float d[4][NP];
    float lf[4];
    lf[0] = adj_qstep<NP, DA, 0>(P, Y, d[0], gv, row, yv, io, t1);
    lf[1] = adj_qstep<NP, DA, 1>(P, Y, d[1], gv, row, yv, io, t1 - 1);
    lf[2] = adj_qstep<NP, DA, 2>(P, Y, d[2], gv, row, yv, io, t1 - 2);
    lf[3] = adj_qstep<NP, DA, 3>(P, Y, d[3], gv, row, yv, io, t1 - 3);
    // refill the slot with the block RB below (times clamped into the signal)
    {
      const int64_t tn = max(min(t1 - 4 * RB - q, io.T - 1), (int64_t)0);
      r.gq[SL] = io.gb[tn];
#pragma unroll
      for (int i = 0; i < NP; ++i) r.aq[SL][i] = io.Ab[tn * NP + i];
      if constexpr (DA) r.yq[SL] = *yaddr(io, t1 - 4 * RB - q - 1 - NP);
    }
    pin();
    // lane q writes its sample's outputs (past T: into the sink)
    const int64_t tm = t1 - q;
    const bool live = tm < io.T;
    if constexpr (DX) *(live ? io.dxb + tm : io.sink) = pick4(q, lf[0], lf[1], lf[2], lf[3]);
    if constexpr (DA) {
      float* dp = live ? io.dAb + tm * NP : io.sink;
#pragma unroll
      for (int i = 0; i < NP; ++i) dp[i] = pick4(q, d[0][i], d[1][i], d[2][i], d[3][i]);
    }
    adj_blocks<NP, DA, DX, SL + 1>(P, Y, r, io, tb, q);
  }
}

template <int NP, bool DA, bool DX>
__device__ __forceinline__ void adj_chunk(const AdjIo& io, double* cr, int64_t t0, int L, int q, float* dz) {
  const int64_t tstart = t0 + L - 1;  // 3 mod 4
  double P[NP];
#pragma unroll
  for (int rr = 0; rr < NP; ++rr) P[NP - 1 - rr] = cr[rr];  // pending sum of position tstart - rr
  float Y[NP];
#pragma unroll
  for (int i = 1; i <= NP; ++i) {
    const int64_t u = tstart - i;
    const float v = *yaddr(io, u);
    Y[(NP - 1 - i + NP) & (NP - 1)] = (DA && yflag(io, u)) ? v : 0.f;
  }
  AdjQ<NP> r;
#pragma unroll
  for (int s = 0; s < RB; ++s) {
    const int64_t t = tstart - 4 * s - q;
    const int64_t tn = max(min(t, io.T - 1), (int64_t)0);
    r.gq[s] = opq(io.gb[tn]);
#pragma unroll
    for (int i = 0; i < NP; ++i) r.aq[s][i] = opq(io.Ab[tn * NP + i]);
    r.yq[s] = DA ? opq(*yaddr(io, t - 1 - NP)) : 0.f;
  }
  settle(P);
  for (int jb = 0; jb < L; jb += 4 * RB) adj_blocks<NP, DA, DX, 0>(P, Y, r, io, tstart - jb, q);
  if (dz && q == 0) {
    // pending sums of positions -1-rr sit in slots NP-1-rr
#pragma unroll
    for (int rr = 0; rr < NP; ++rr)
      if (rr < io.n) dz[rr] = (float)P[NP - 1 - rr];
  }
}

template <int NP, bool DA>
__global__ __launch_bounds__(THREADS) void quad_bwd_rerun(ChunkArgs<float> a) {
  const int64_t gt = (int64_t)blockIdx.x * THREADS + threadIdx.x;
  const int64_t cid = gt >> 2;
  const int q = (int)(gt & 3);
  if (cid >= (a.rg ? a.G : a.B * a.P)) return;  // a quad's lanes leave together
  const int64_t b = chunk_signal(a, cid);
  const Span sp = sig_span(a, b);
  const int64_t c = cid - sp.cbase;
  double* cr = a.carry + cid * NP;
  AdjIo io;
  io.Ab = a.A + sp.ao;
  io.gb = a.x + sp.xo;  // upstream gradient
  io.yb = a.y + sp.xo;
  io.haszi = a.zi != nullptr;
  io.zb = a.zi ? a.zi + sp.zo : io.yb;
  io.dxb = a.dx ? a.dx + sp.xo : nullptr;
  io.dAb = DA ? a.dA + sp.dao : nullptr;
  // the carry is read (by every lane of the quad) before any store: its NP
  // doubles hold a dx value and a dA row
  io.sink = reinterpret_cast<float*>(cr);
  io.T = sp.T;
  io.n = a.n;
  float* dz = (c == 0 && a.dzi) ? a.dzi + sp.zo : nullptr;
  if (io.dxb)
    adj_chunk<NP, DA, true>(io, cr, c * a.L, a.L, q, dz);
  else
    adj_chunk<NP, DA, false>(io, cr, c * a.L, a.L, q, dz);
}

inline dim3 quad_grid(const ChunkArgs<float>& a) {
  const int64_t lanes = (a.rg ? a.G : a.B * a.P) * 4;
  return dim3((unsigned)((lanes + THREADS - 1) / THREADS));
}


// ---------------------------------------------------------------------------
// Chunk summaries at orders 2, 4 and 8 (phase 1), one quad per chunk with the
// same sample layout (lane q loads sample 4k + q of block k; each step
// broadcasts its sample to the quad). Lane q owns the NB = max(NP / 4, 1) basis
// indices q + 4m (at order 2 lanes 2 and 3 repeat lanes 0 and 1 and store
// nothing). Layouts as in the other summaries: out[k * NP + r] = Phi[k][r]
// (Psi[k][r]), out[NP * NP + k] = z[k] (w[k]).
//   forward  descending, the transposed recurrence from a unit at output k
//            (sample L-1-k): the pending sums left before the chunk are row k
//            of Phi, and z[k] = sum_t lambda_t x_t. Forward summaries cover whole
//            chunks.
//   adjoint  ascending, the zero-input recurrence from a unit at the state
//            entering the chunk (position -1-r): its final window is row r of
//            Psi, and w[r] = sum_t g_t h_r(t); samples past the end of the
//            signal have zero rows and g (selected at use).
template <int NP>
struct SumQ {
  static constexpr int NB = NP >= 4 ? NP / 4 : 1;  // basis vectors per lane
  float vq[RB];      // x (forward) or g (adjoint) of the lane's sample
  float aq[RB][NP];  // its row
};

template <int NP>
__device__ __forceinline__ int basis_of(int q, int m) {
  return NP >= 4 ? q + 4 * m : (q & (NP - 1));
}

// Step J of block SL of a pass: a pass starts at L-1-16m (7 mod 8: chunk lengths
// are multiples of 16), so its block SL starts at 7 - 4 SL mod 8.
template <int NP, int SL, int J>
__device__ __forceinline__ void fsum_qstep(double (&p)[SumQ<NP>::NB][NP], double (&z)[SumQ<NP>::NB], float xv,
                                           const float (&row)[NP]) {
  constexpr int S = (7 - 4 * SL - J) & (NP - 1);  // slot of the step's position
  double c[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) c[i] = (double)bq<J>(row[i]);
  const double xd = (double)bq<J>(xv);
#pragma unroll
  for (int m = 0; m < SumQ<NP>::NB; ++m) {
    const double lam = p[m][S];
    p[m][S] = 0.0;
#pragma unroll
    for (int i = 1; i <= NP; ++i) p[m][(S - i) & (NP - 1)] = fma(-c[i - 1], lam, p[m][(S - i) & (NP - 1)]);
    z[m] = fma(lam, xd, z[m]);
  }
}

template <int NP, int SL>
__device__ __forceinline__ void fsum_blocks(double (&p)[SumQ<NP>::NB][NP], double (&z)[SumQ<NP>::NB],
                                            SumQ<NP>& r, const float* Ab, const float* xb, int tb, int q) {
  if constexpr (SL < RB) {
    const int t1 = tb - 4 * SL;  // samples t1 - j, j = 0..3; lane q's sample t1 - q
    float row[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) row[i] = r.aq[SL][i];
    const float xv = r.vq[SL];
    fsum_qstep<NP, SL, 0>(p, z, xv, row);
    fsum_qstep<NP, SL, 1>(p, z, xv, row);
    fsum_qstep<NP, SL, 2>(p, z, xv, row);
    fsum_qstep<NP, SL, 3>(p, z, xv, row);
    {
      const int tn = max(t1 - 4 * RB - q, 0);
      r.vq[SL] = xb[tn];
#pragma unroll
      for (int i = 0; i < NP; ++i) r.aq[SL][i] = Ab[(int64_t)tn * NP + i];
    }
    pin();
    fsum_blocks<NP, SL + 1>(p, z, r, Ab, xb, tb, q);
  }
}

template <int NP>
__global__ __launch_bounds__(THREADS) void quad_fwd_summary(ChunkArgs<float> a) {
  constexpr int NB = SumQ<NP>::NB;
  const int64_t gt = (int64_t)blockIdx.x * THREADS + threadIdx.x;
  const int64_t item = gt >> 2;
  const int q = (int)(gt & 3);
  SumItem it;
  if (!summary_item<NP, true>(a, item, it)) return;  // a quad's lanes leave together
  const int L = a.L;
  double p[NB][NP], z[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int k = basis_of<NP>(q, m);
#pragma unroll
    for (int i = 0; i < NP; ++i) p[m][i] = (i == NP - 1 - k) ? 1.0 : 0.0;  // unit at sample L-1-k
    z[m] = 0.0;
  }
  SumQ<NP> r;
#pragma unroll
  for (int s = 0; s < RB; ++s) {
    const int tn = max(L - 1 - 4 * s - q, 0);
    r.vq[s] = opq(it.xb[tn]);
#pragma unroll
    for (int i = 0; i < NP; ++i) r.aq[s][i] = opq(it.Ab[(int64_t)tn * NP + i]);
  }
  for (int jb = 0; jb < L; jb += 4 * RB) fsum_blocks<NP, 0>(p, z, r, it.Ab, it.xb, L - 1 - jb, q);
  if (q < NP) {
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int k = basis_of<NP>(q, m);
      // the pending sum of position -1-rr sits in slot NP-1-rr
#pragma unroll
      for (int rr = 0; rr < NP; ++rr) it.out[k * NP + rr] = p[m][NP - 1 - rr];
      it.out[NP * NP + k] = z[m];
    }
  }
}

// Step J of block SL of a pass: a pass starts at 16m, so its block SL starts at
// 4 SL mod 8.
template <int NP, int SL, int J>
__device__ __forceinline__ void asum_qstep(double (&h)[SumQ<NP>::NB][NP], double (&w)[SumQ<NP>::NB], float gv,
                                           const float (&row)[NP], bool live) {
  constexpr int S = (4 * SL + J) & (NP - 1);  // slot of the step's position
  double c[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const float v = bq<J>(row[i]);
    c[i] = (double)(live ? v : 0.f);
  }
  const float gf = bq<J>(gv);
  const double g = (double)(live ? gf : 0.f);
#pragma unroll
  for (int m = 0; m < SumQ<NP>::NB; ++m) {
    double acc = 0.0;
#pragma unroll
    for (int i = NP; i >= 1; --i) acc = fma(-c[i - 1], h[m][(S - i) & (NP - 1)], acc);
    h[m][S] = acc;
    w[m] = fma(g, acc, w[m]);
  }
}

template <int NP, int SL>
__device__ __forceinline__ void asum_blocks(double (&h)[SumQ<NP>::NB][NP], double (&w)[SumQ<NP>::NB],
                                            SumQ<NP>& r, const float* Ab, const float* gb, int live, int tb,
                                            int q) {
  if constexpr (SL < RB) {
    const int t1 = tb + 4 * SL;  // samples t1 + j; lane q's sample t1 + q
    float row[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) row[i] = r.aq[SL][i];
    const float gv = r.vq[SL];
    asum_qstep<NP, SL, 0>(h, w, gv, row, t1 < live);
    asum_qstep<NP, SL, 1>(h, w, gv, row, t1 + 1 < live);
    asum_qstep<NP, SL, 2>(h, w, gv, row, t1 + 2 < live);
    asum_qstep<NP, SL, 3>(h, w, gv, row, t1 + 3 < live);
    {
      const int tn = min(t1 + 4 * RB + q, live - 1);
      r.vq[SL] = gb[tn];
#pragma unroll
      for (int i = 0; i < NP; ++i) r.aq[SL][i] = Ab[(int64_t)tn * NP + i];
    }
    pin();
    asum_blocks<NP, SL + 1>(h, w, r, Ab, gb, live, tb, q);
  }
}

template <int NP>
__global__ __launch_bounds__(THREADS) void quad_adj_summary(ChunkArgs<float> a) {
  constexpr int NB = SumQ<NP>::NB;
  const int64_t gt = (int64_t)blockIdx.x * THREADS + threadIdx.x;
  const int64_t item = gt >> 2;
  const int q = (int)(gt & 3);
  SumItem it;
  if (!summary_item<NP, false>(a, item, it)) return;  // a quad's lanes leave together
  const int L = a.L, live = it.live;
  double h[NB][NP], w[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int rr = basis_of<NP>(q, m);
#pragma unroll
    for (int i = 0; i < NP; ++i) h[m][i] = (i == NP - 1 - rr) ? 1.0 : 0.0;  // unit at position -1-rr
    w[m] = 0.0;
  }
  SumQ<NP> r;
#pragma unroll
  for (int s = 0; s < RB; ++s) {
    const int tn = min(4 * s + q, live - 1);
    r.vq[s] = opq(it.xb[tn]);  // the upstream gradient
#pragma unroll
    for (int i = 0; i < NP; ++i) r.aq[s][i] = opq(it.Ab[(int64_t)tn * NP + i]);
  }
  for (int jb = 0; jb < L; jb += 4 * RB) asum_blocks<NP, 0>(h, w, r, it.Ab, it.xb, live, jb, q);
  if (q < NP) {
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int rr = basis_of<NP>(q, m);
      // Psi[rr][kk] = h_rr(L-1-kk), in slot NP-1-kk
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) it.out[rr * NP + kk] = h[m][NP - 1 - kk];
      it.out[NP * NP + rr] = w[m];
    }
  }
}

inline dim3 quad_sum_grid(const ChunkArgs<float>& a) {
  const int64_t items = a.rg ? a.G : a.B * (a.P - 1);
  return dim3((unsigned)((items * 4 + THREADS - 1) / THREADS));
}

}  // namespace quadr

int launch_quad_summary(ChunkArgs<float> a, int np, bool fwd, hipStream_t st) {
  const dim3 g = quadr::quad_sum_grid(a);
  if (g.x == 0) return 0;
  const dim3 b(quadr::THREADS);
  switch (np) {
    case 2:
      if (fwd) hipLaunchKernelGGL(quadr::quad_fwd_summary<2>, g, b, 0, st, a);
      else hipLaunchKernelGGL(quadr::quad_adj_summary<2>, g, b, 0, st, a);
      break;
    case 4:
      if (fwd) hipLaunchKernelGGL(quadr::quad_fwd_summary<4>, g, b, 0, st, a);
      else hipLaunchKernelGGL(quadr::quad_adj_summary<4>, g, b, 0, st, a);
      break;
    case 8:
      if (fwd) hipLaunchKernelGGL(quadr::quad_fwd_summary<8>, g, b, 0, st, a);
      else hipLaunchKernelGGL(quadr::quad_adj_summary<8>, g, b, 0, st, a);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_quad_fwd_rerun(ChunkArgs<float> a, int np, hipStream_t st) {
  const dim3 g = quadr::quad_grid(a);
  if (g.x == 0) return 0;
  switch (np) {
    case 2: hipLaunchKernelGGL(quadr::quad_fwd_rerun<2>, g, dim3(quadr::THREADS), 0, st, a); break;
    case 4: hipLaunchKernelGGL(quadr::quad_fwd_rerun<4>, g, dim3(quadr::THREADS), 0, st, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_quad_bwd_rerun(ChunkArgs<float> a, int np, hipStream_t st) {
  const dim3 g = quadr::quad_grid(a);
  if (g.x == 0) return 0;
  const bool da = a.dA != nullptr;
  switch (np) {
    case 2:
      if (da) hipLaunchKernelGGL((quadr::quad_bwd_rerun<2, true>), g, dim3(quadr::THREADS), 0, st, a);
      else hipLaunchKernelGGL((quadr::quad_bwd_rerun<2, false>), g, dim3(quadr::THREADS), 0, st, a);
      break;
    case 4:
      if (da) hipLaunchKernelGGL((quadr::quad_bwd_rerun<4, true>), g, dim3(quadr::THREADS), 0, st, a);
      else hipLaunchKernelGGL((quadr::quad_bwd_rerun<4, false>), g, dim3(quadr::THREADS), 0, st, a);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // namespace tlpc
