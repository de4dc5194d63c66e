// Chunked scan for order 16, float32 data (the B=256, T=48000, N=16 shape and
// any other float32 order-16 batch), built from the DPP blocks of tlpc_dpp16.h:
//
//   summary  16 chunks per workgroup: waves 0-3 are 16 matrix rows (one chunk
//            per 16-lane row, float64 basis responses -> Phi), wave 4 is 16 quads
//            (float64 zero-state response -> z). Forward: chunks 0..P-2; adjoint
//            (transposed recurrence, reverse time): chunks 1..P-1.
//   carry    the generic per-signal chain (tlpc_chunked_impl.h), float64.
//   rerun    one quad per chunk from its float64 carry, float64 arithmetic on the
//            float32 data (the rerun produces the outputs, so it keeps their
//            error at the final float32 rounding);
//            the adjoint rerun emits dx, the 64-byte dA row of every sample from
//            the quad's four lanes (one float4 each, a coalesced row store) and
//            dzi for chunk 0.
//
// Summaries and carries use the generic layout: summ[chunk][r*16 + k] = Phi[r][k],
// summ[chunk][256 + r] = z[r]; carry[b][c][r].
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"

namespace tlpc {
namespace c16 {

using namespace dpp16;

constexpr int SUM_THREADS = 320;  // 4 matrix waves + 1 quad wave = 16 chunks
constexpr int RERUN_THREADS = 256;

// ---------------------------------------------------------------------------
// Matrix rows.
template <int J>
__device__ __forceinline__ void fwd_row_block(double (&r)[NP], float (&cf)[NP], const float* ap,
                                              int jb, int L) {
  if constexpr (J < NP) {
    const int jn = jb + J + 8;
    cf[(J + 8) & 15] = (jn < L) ? ap[(int64_t)jn * NP] : 0.f;
    r[J] = fwd_row_step<J>((double)cf[J], r);
    fwd_row_block<J + 1>(r, cf, ap, jb, L);
  }
}

// Descending steps jb+15 .. jb; step jb+JJ reads cf[JJ] and loads step jb+JJ-8.
template <int JJ>
__device__ __forceinline__ void adj_row_block(double (&p)[NP], float (&cf)[NP], const float* ap,
                                              int jb, int64_t t0, int64_t T) {
  if constexpr (JJ >= 0) {
    const int jn = jb + JJ - 8;
    const float a = cf[JJ];
    cf[(JJ + 8) & 15] = (jn >= 0 && t0 + jn < T) ? ap[(int64_t)jn * NP] : 0.f;
    adj_row_step<JJ>((double)a, p);
    adj_row_block<JJ - 1>(p, cf, ap, jb, t0, T);
  }
}

// ---------------------------------------------------------------------------
// Quad runs (prefetching PF rows ahead).
template <int JJ, bool STORE, typename V>
__device__ __forceinline__ void fwd_quad_block(V (&R)[4], float4 (&cq)[PF], float (&xq)[PF],
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
      const V y = fwd_quad_step<JJ & 3>(R, c, xv, ql);
      if (STORE && ((t & 3) == ql)) yb[t] = (float)y;
    }
    fwd_quad_block<JJ + 1, STORE>(R, cq, xq, Ab, xb, yb, t0, jb, len, ql);
  }
}

template <bool STORE, typename V>
__device__ __forceinline__ void fwd_quad_run(V (&R)[4], const float* Ab, const float* xb,
                                             float* yb, int64_t t0, int len, int ql) {
  float4 cq[PF];
  float xq[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    if (j < len) {
      cq[j] = reinterpret_cast<const float4*>(Ab + (t0 + j) * NP)[ql];
      xq[j] = xb[t0 + j];
    }
  }
  for (int jb = 0; jb < len; jb += PF) fwd_quad_block<0, STORE>(R, cq, xq, Ab, xb, yb, t0, jb, len, ql);
}

// ---------------------------------------------------------------------------
// Forward summaries: chunks c = 0..P-2 (full length L).
__global__ __launch_bounds__(SUM_THREADS) void sum_fwd(ChunkArgs<float> a) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t cps = a.P - 1;
  const int64_t nch = a.B * cps;
  const int L = a.L;
  if (wave < 4) {
    const int64_t cid = (int64_t)blockIdx.x * 16 + wave * 4 + (lane >> 4);
    if (cid >= nch) return;
    const int kk = lane & 15;
    const int64_t b = cid / cps;
    const int64_t c = cid - b * cps;
    const float* ap = a.A + (b * a.T + c * L) * NP + kk;
    double r[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) r[i] = (i == NP - 1 - kk) ? 1.0 : 0.0;  // y_{-1-kk} = 1
    float cf[NP];
#pragma unroll
    for (int j = 0; j < 8; ++j) cf[j] = ap[(int64_t)j * NP];
    for (int jb = 0; jb < L; jb += NP) fwd_row_block<0>(r, cf, ap, jb, L);
    double* out = a.summ + cid * (NP * NP + NP);
#pragma unroll
    for (int rr = 0; rr < NP; ++rr) out[rr * NP + kk] = r[NP - 1 - rr];
  } else {
    const int64_t cid = (int64_t)blockIdx.x * 16 + (lane >> 2);
    if (cid >= nch) return;
    const int ql = lane & 3;
    const int64_t b = cid / cps;
    const int64_t c = cid - b * cps;
    double R[4] = {0.0, 0.0, 0.0, 0.0};
    fwd_quad_run<false>(R, a.A + b * a.T * NP, a.x + b * a.T, nullptr, c * L, L, ql);
    double* out = a.summ + cid * (NP * NP + NP) + NP * NP;
#pragma unroll
    for (int m = 0; m < 4; ++m) out[4 * ql + m] = R[3 - m];  // z[4ql+m] = y_{L-1-4ql-m}
  }
}

// Adjoint summaries: chunks c = 1..P-1 (the last may be partial; its steps past
// the end are no-ops of the zero state, and its Psi is never used).
__global__ __launch_bounds__(SUM_THREADS) void sum_adj(ChunkArgs<float> a) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t cps = a.P - 1;
  const int64_t nch = a.B * cps;
  const int L = a.L;
  const int64_t T = a.T;
  if (wave < 4) {
    const int64_t cid = (int64_t)blockIdx.x * 16 + wave * 4 + (lane >> 4);
    if (cid >= nch) return;
    const int kk = lane & 15;
    const int64_t b = cid / cps;
    const int64_t c = cid - b * cps + 1;
    const int64_t t0 = c * L;
    const float* ap = a.A + (b * T + t0) * NP + kk;
    double p[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) p[i] = (i == NP - 1 - kk) ? 1.0 : 0.0;  // pending(t1-1-kk) = 1
    float cf[NP];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int s = L - 1 - j;
      cf[15 - j] = (t0 + s < T) ? ap[(int64_t)s * NP] : 0.f;
    }
    for (int jb = L - NP; jb >= 0; jb -= NP) adj_row_block<15>(p, cf, ap, jb, t0, T);
    double* out = a.summ + cid * (NP * NP + NP);
#pragma unroll
    for (int rr = 0; rr < NP; ++rr) out[rr * NP + kk] = p[NP - 1 - rr];
  } else {
    const int64_t cid = (int64_t)blockIdx.x * 16 + (lane >> 2);
    if (cid >= nch) return;
    const int ql = lane & 3;
    const int64_t b = cid / cps;
    const int64_t c = cid - b * cps + 1;
    const int64_t t0 = c * L;
    const float* Ab = a.A + b * T * NP;
    const float* gb = a.x + b * T;
    double P[4] = {0.0, 0.0, 0.0, 0.0};
    double now = 0.0;
    // descending steps L-1 .. 0, PF rows in flight
    float4 cq[PF];
    float gq[PF];
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int64_t t = t0 + L - 1 - j;
      cq[j] = (t < T) ? reinterpret_cast<const float4*>(Ab + t * NP)[ql] : zero4;
      gq[j] = (t < T) ? gb[t] : 0.f;
    }
    for (int jb = 0; jb < L; jb += PF) {
      // step index from the top: s = jb + jj, t = t0 + L - 1 - s, slot (3 - jj) & 3
#define TLPC_ADJ_Q(JJ)                                                                 \
  {                                                                                    \
    const int s = jb + JJ;                                                             \
    const float4 cv = cq[JJ];                                                          \
    const float gv = gq[JJ];                                                           \
    const int64_t tn = t0 + L - 1 - (s + PF);                                          \
    if (s + PF < L) {                                                                  \
      cq[JJ] = (tn < T) ? reinterpret_cast<const float4*>(Ab + tn * NP)[ql] : zero4;   \
      gq[JJ] = (tn < T) ? gb[tn] : 0.f;                                                \
    }                                                                                  \
    adj_quad_step<(3 - JJ) & 3>(P, now, cv, gv, ql);                                   \
  }
      TLPC_ADJ_Q(0) TLPC_ADJ_Q(1) TLPC_ADJ_Q(2) TLPC_ADJ_Q(3)
      TLPC_ADJ_Q(4) TLPC_ADJ_Q(5) TLPC_ADJ_Q(6) TLPC_ADJ_Q(7)
#undef TLPC_ADJ_Q
    }
    // w[0] = now (lane 0); w[4ql+1+m] = P[(2-m) & 3]
    double* out = a.summ + cid * (NP * NP + NP) + NP * NP;
    if (ql == 0) out[0] = now;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int rr = 4 * ql + 1 + m;
      if (rr < NP) out[rr] = P[(2 - m) & 3];
    }
  }
}

// ---------------------------------------------------------------------------
// Forward rerun: one quad per chunk (all P chunks of all signals).
__global__ __launch_bounds__(RERUN_THREADS) void rerun_fwd(ChunkArgs<float> a) {
  const int64_t cid = ((int64_t)blockIdx.x * RERUN_THREADS + threadIdx.x) >> 2;
  if (cid >= a.B * a.P) return;
  const int ql = threadIdx.x & 3;
  const int64_t b = cid / a.P;
  const int64_t c = cid - b * a.P;
  const int64_t t0 = c * a.L;
  const int len = (int)((a.T - t0) < a.L ? (a.T - t0) : a.L);
  const double* cr = a.carry + cid * NP;
  double R[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) R[3 - m] = cr[4 * ql + m];  // y_{t0-1-4ql-m}, slot 3-m
  fwd_quad_run<true>(R, a.A + b * a.T * NP, a.x + b * a.T, a.y + b * a.T, t0, len, ql);
}

// Adjoint rerun: one quad per chunk, descending from the chunk's last step (steps
// past the end of the last chunk are no-ops of its zero carry).
__global__ __launch_bounds__(RERUN_THREADS) void rerun_adj(ChunkArgs<float> a) {
  const int64_t cid = ((int64_t)blockIdx.x * RERUN_THREADS + threadIdx.x) >> 2;
  if (cid >= a.B * a.P) return;
  const int ql = threadIdx.x & 3;
  const int64_t b = cid / a.P;
  const int64_t c = cid - b * a.P;
  const int64_t T = a.T;
  const int L = a.L;
  const int64_t t0 = c * L;
  const float* Ab = a.A + b * T * NP;
  const float* gb = a.x + b * T;
  const float* yb = a.y + b * T;
  const float* zib = a.zi ? a.zi + b * NP : nullptr;
  float* dxb = a.dx ? a.dx + b * T : nullptr;
  float* dAb = a.dA ? a.dA + b * T * NP : nullptr;
  const bool need_y = dAb != nullptr;
  // pending state from the carry: now = p_0, window slot (2-m)&3 = p_{4ql+1+m}
  const double* cr = a.carry + cid * NP;
  double P[4];
  double now = cr[0];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int rr = 4 * ql + 1 + m;
    P[(2 - m) & 3] = rr < NP ? cr[rr] : 0.0;
  }
  auto yval = [&](int64_t u) -> float {
    if (u >= 0) return u < T ? yb[u] : 0.f;
    return (zib && -u - 1 < NP) ? zib[-u - 1] : 0.f;
  };
  // y window: slot u mod 4 holds y_u, u in [t-4-4ql, t-1-4ql], t = t0 + L - 1
  float Y[4] = {0.f, 0.f, 0.f, 0.f};
  const int64_t tstart = t0 + L - 1;
  if (need_y) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int64_t u = tstart - 1 - 4 * ql - m;
      Y[(2 - m) & 3] = yval(u);  // (tstart - 1 - m) mod 4 = (2 - m) & 3
    }
  }
  float4 cq[PF];
  float gq[PF], yq[PF];
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int64_t t = tstart - j;
    cq[j] = (t < T) ? reinterpret_cast<const float4*>(Ab + t * NP)[ql] : zero4;
    gq[j] = (t < T) ? gb[t] : 0.f;
    yq[j] = need_y ? yval(t - 17) : 0.f;  // enters lane 3's window after step t
  }
  for (int jb = 0; jb < L; jb += PF) {
#define TLPC_ADJ_R(JJ)                                                                        \
  {                                                                                           \
    constexpr int J = (3 - JJ) & 3;                                                           \
    const int s = jb + JJ;                                                                    \
    const int64_t t = tstart - s;                                                             \
    const float4 cv = cq[JJ];                                                                 \
    const float gv = gq[JJ];                                                                  \
    const float ynew = yq[JJ];                                                                \
    const int64_t tn = t - PF;                                                                \
    if (s + PF < L) {                                                                         \
      cq[JJ] = (tn < T) ? reinterpret_cast<const float4*>(Ab + tn * NP)[ql] : zero4;          \
      gq[JJ] = (tn < T) ? gb[tn] : 0.f;                                                       \
      yq[JJ] = need_y ? yval(tn - 17) : 0.f;                                                  \
    }                                                                                         \
    const double lam = adj_quad_step<J>(P, now, cv, gv, ql);                                  \
    if (t < T) {                                                                              \
      if (dxb && (t & 3) == ql) dxb[t] = (float)lam;                                          \
      if (need_y) {                                                                           \
        float4 d;                                                                             \
        d.x = (float)(-lam * Y[(J - 1) & 3]);                                                 \
        d.y = (float)(-lam * Y[(J - 2) & 3]);                                                 \
        d.z = (float)(-lam * Y[(J - 3) & 3]);                                                 \
        d.w = (float)(-lam * Y[(J - 4) & 3]);                                                 \
        reinterpret_cast<float4*>(dAb + t * NP)[ql] = d;                                      \
      }                                                                                       \
    }                                                                                         \
    if (need_y) {                                                                             \
      constexpr int S = (J - 1) & 3;                                                          \
      const float down = qperm<QP_DOWN>(Y[S]);                                                \
      Y[S] = (ql == 3) ? ynew : down;                                                         \
    }                                                                                         \
  }
    TLPC_ADJ_R(0) TLPC_ADJ_R(1) TLPC_ADJ_R(2) TLPC_ADJ_R(3)
    TLPC_ADJ_R(4) TLPC_ADJ_R(5) TLPC_ADJ_R(6) TLPC_ADJ_R(7)
#undef TLPC_ADJ_R
  }
  if (c == 0 && a.dzi) {
    // pending of positions -1-r: r = 0 in `now` (lane 0), r = 4ql+1+m in slot (2-m)&3
    float* dz = a.dzi + b * NP;
    if (ql == 0) dz[0] = (float)now;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int rr = 4 * ql + 1 + m;
      if (rr < NP) dz[rr] = (float)P[(2 - m) & 3];
    }
  }
}

// ---------------------------------------------------------------------------
// Carry chain, one wave per signal. Summaries come through LDS in batches of
// CB links: the loads of batch n+1 are in flight while batch n is chained.
constexpr int CB = 8;
constexpr int SUMM = NP * NP + NP;

template <bool FWD>
__global__ __launch_bounds__(64) void carry16(ChunkArgs<float> a) {
  __shared__ double buf[2][CB][SUMM];
  const int64_t b = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t P = a.P;
  const int64_t links = P - 1;
  const double* summ = a.summ + b * links * SUMM;
  double* carry = a.carry + b * P * NP;
  // link l maps the carry of step l to step l+1; forward: chunk l, summary l;
  // adjoint: chunk P-1-l, summary P-2-l
  auto src = [&](int64_t l) -> const double* { return summ + (FWD ? l : links - 1 - l) * SUMM; };
  constexpr int PER = CB * SUMM / 64;  // doubles per lane per batch
  double stage[PER];
  auto fetch = [&](int64_t l0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = q * 64 + lane;
      const int li = e / SUMM, off = e - li * SUMM;
      stage[q] = (l0 + li < links) ? src(l0 + li)[off] : 0.0;
    }
  };
  auto commit = [&](int bb) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = q * 64 + lane;
      buf[bb][e / SUMM][e % SUMM] = stage[q];
    }
  };
  double s = 0.0;
  if (FWD && lane < NP && a.zi) s = (double)a.zi[b * NP + lane];
  const int row = lane & 15;
  if (links > 0) {
    fetch(0);
    commit(0);
  }
  for (int64_t l0 = 0, bb = 0; l0 < links; l0 += CB, bb ^= 1) {
    if (l0 + CB < links) fetch(l0 + CB);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this batch is in LDS
    const int nb = (int)((links - l0) < CB ? (links - l0) : CB);
    for (int li = 0; li < nb; ++li) {
      const int64_t l = l0 + li;
      if (lane < NP) carry[(FWD ? l : P - 1 - l) * NP + lane] = s;
      double m[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) m[k] = buf[bb][li][row * NP + k];
      s = chain_link(s, m, buf[bb][li][NP * NP + row]);
    }
    if (l0 + CB < links) commit(bb ^ 1);
  }
  if (lane < NP) carry[(FWD ? P - 1 : 0) * NP + lane] = s;
}

}  // namespace c16

int launch_chunk16_carry(ChunkArgs<float> a, bool fwd, hipStream_t st) {
  if (fwd)
    hipLaunchKernelGGL(c16::carry16<true>, dim3((unsigned)a.B), dim3(64), 0, st, a);
  else
    hipLaunchKernelGGL(c16::carry16<false>, dim3((unsigned)a.B), dim3(64), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_chunk16_fwd_summary(ChunkArgs<float> a, hipStream_t st) {
  const int64_t n = a.B * (a.P - 1);
  if (n > 0)
    hipLaunchKernelGGL(c16::sum_fwd, dim3((unsigned)((n + 15) / 16)), dim3(c16::SUM_THREADS), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}
int launch_chunk16_adj_summary(ChunkArgs<float> a, hipStream_t st) {
  const int64_t n = a.B * (a.P - 1);
  if (n > 0)
    hipLaunchKernelGGL(c16::sum_adj, dim3((unsigned)((n + 15) / 16)), dim3(c16::SUM_THREADS), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}
int launch_chunk16_fwd_rerun(ChunkArgs<float> a, hipStream_t st) {
  const int64_t lanes = a.B * a.P * 4;
  hipLaunchKernelGGL(c16::rerun_fwd, dim3((unsigned)((lanes + c16::RERUN_THREADS - 1) / c16::RERUN_THREADS)),
                     dim3(c16::RERUN_THREADS), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}
int launch_chunk16_adj_rerun(ChunkArgs<float> a, hipStream_t st) {
  const int64_t lanes = a.B * a.P * 4;
  hipLaunchKernelGGL(c16::rerun_adj, dim3((unsigned)((lanes + c16::RERUN_THREADS - 1) / c16::RERUN_THREADS)),
                     dim3(c16::RERUN_THREADS), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // namespace tlpc
