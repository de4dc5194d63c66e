// Chunked linear-recurrence scan for the all-pole filter (any batch size).
//
// Forward recurrence (reference: torchlpc/csrc/scan_cpu.cpp:112-120):
//   y[t] = x[t] - sum_{i=1..N} A[t,i] y[t-i]
// The state s_t = [y_{t-1}, ..., y_{t-N}] evolves by a companion matrix, so a chunk
// of L samples is the affine map s_out = Phi s_in + z. Three phases:
//   1. summary: per chunk, Phi and z, computed in the wide type by NP (forward:
//      an adjoint sweep per row of Phi) or NP+1 (adjoint) lanes that share the
//      chunk's A rows from LDS; at order 32 the adjoint summary runs in forward
//      form with NP lanes;
//   2. carry:   per signal, s_{c+1} = Phi_c s_c + z_c from s_0 = zi;
//   3. rerun:   per chunk, one lane re-filters its L samples from the carry and
//               writes y.
// The adjoint (reference backward: torchlpc/core.py:186-229) runs the same three
// phases in reverse time on the transposed recurrence. Its state is the vector of
// pending sums p_j for positions t-1-j; one step is
//   lambda[t] = g[t] + p(t);  p(t-j) -= conj(A[t,j]) * lambda[t]   (j = 1..N)
// and the rerun phase also emits dx = lambda, dA[t,i] = -lambda[t] conj(y[t-i])
// (reference: core.py:219-227) and dzi (core.py:215-216).
//
// NP is the compiled order (a power of two >= the runtime order n); coefficients
// for lags n < i <= NP are zero, so padded lags never change a result. The state
// ring is indexed statically: time t lives in slot t mod NP, every chunk starts
// at a multiple of NP, and the unrolled inner loop covers NP steps.
#pragma once

#include <type_traits>

#include "tlpc_common.h"
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"

namespace tlpc {

// LANES: lanes per summary item (the forward summary uses NP, the adjoint NP+1).
template <typename S, int NP, int LANES = NP + 1>
struct ChunkCfg {
  static constexpr int TB = NP < 16 ? 16 : NP;  // time steps staged per LDS block
  static constexpr int LDS_BUDGET = 40 * 1024;
  static constexpr int G_LANES = 256 / LANES;
  using W = typename Wide<S>::type;  // summaries and carries
  static constexpr int G_LDS = LDS_BUDGET / (TB * (NP + 1) * (int)sizeof(W));
  static constexpr int G = (G_LANES < G_LDS ? G_LANES : G_LDS) < 1 ? 1 : (G_LANES < G_LDS ? G_LANES : G_LDS);
  static constexpr int THREADS = ((G * LANES + 63) / 64) * 64;
  static constexpr int ASTRIDE = TB * NP + (16 / (int)sizeof(W) > 0 ? 16 / (int)sizeof(W) : 1);
  static constexpr int SUMM = NP * NP + NP;  // elements per chunk summary
};

// ----------------------------------------------------------------------------
// Summary work items. Item i of a uniform batch is chunk (i / (P-1), c) with
// c = i % (P-1) (+1 for the adjoint); in a ragged batch item i is chunk id i,
// and the chunks without a summary (the last of a signal forward, the first
// adjoint) are idle. Per item the workgroup keeps, in LDS, where the chunk's A
// rows and inputs start and how many of its samples exist.
template <typename S, int NP, int LANES = NP + 1>
struct ItemTab {
  static constexpr int G = ChunkCfg<S, NP, LANES>::G;
  int64_t a0[G];   // offset of A at the chunk's first sample
  int64_t x0[G];   // offset of the chunk's first input sample
  int64_t rem[G];  // samples of the signal from the chunk start (<= 0: idle)
};

template <typename S, int NP, int LANES>
__device__ __forceinline__ void fill_items(const ChunkArgs<S>& a, ItemTab<S, NP, LANES>& tab, int64_t first,
                                           bool fwd) {
  using C = ChunkCfg<S, NP, LANES>;
  const int gg = threadIdx.x;
  if (gg >= C::G) return;
  const int64_t item = first + gg;
  int64_t k = 0, c = 0;
  bool ok;
  if (a.rg) {
    ok = item < a.G;
    if (ok) {
      k = chunk_signal(a, item);
      c = item - sig_span(a, k).cbase;
    }
  } else {
    const int64_t cps = a.P - 1;
    ok = item < a.B * cps;
    if (ok) {
      k = item / cps;
      c = item - k * cps + (fwd ? 0 : 1);
    }
  }
  tab.rem[gg] = -1;
  if (!ok) return;
  const Span sp = sig_span(a, k);
  if (fwd ? (c >= sp.P - 1) : (c < 1)) return;
  const int64_t t = c * a.L;
  tab.a0[gg] = sp.ao + t * sp.at;
  tab.x0[gg] = sp.xo + t;
  tab.rem[gg] = sp.T - t;
}

// Staging: a thread stages the same elements (item, step in the block, lag) in
// every time block of its items, so their addresses are worked out once
// (stage_map): the offset of each element in the block at t0 = 0, and `lim`,
// the element being inside its signal (and its lag below n) while t0 < lim.
template <typename S, int NP, int LANES>
struct StageMap {
  using C = ChunkCfg<S, NP, LANES>;
  static constexpr int NT = C::THREADS;
  static constexpr int NA = C::G * C::TB * NP;
  static constexpr int NX = C::G * C::TB;
  static constexpr int ITA = (NA + NT - 1) / NT;
  static constexpr int ITX = (NX + NT - 1) / NT;
  int64_t pa[ITA], px[ITX];
  int lima[ITA], limx[ITX];
};

template <typename S, int NP, int LANES>
__device__ __forceinline__ void stage_map(const ChunkArgs<S>& a, StageMap<S, NP, LANES>& m,
                                          const ItemTab<S, NP, LANES>& tab) {
  using C = ChunkCfg<S, NP, LANES>;
  using SM = StageMap<S, NP, LANES>;
  constexpr int NT = SM::NT, NA = SM::NA, NX = SM::NX;
  const int64_t at = a.rg ? a.n : a.a_tstride;  // A's time stride, the same for every signal
#pragma unroll
  for (int it = 0; it < SM::ITA; ++it) {
    const int idx = it * NT + (int)threadIdx.x;
    const int gg = min(idx, NA - 1) / (C::TB * NP);
    const int r = min(idx, NA - 1) - gg * (C::TB * NP);
    const int tb = r / NP;
    const int i = r - tb * NP;
    const int64_t rem = tab.rem[gg];
    const bool ok = idx < NA && i < a.n && rem > tb;
    m.lima[it] = ok ? (int)min(rem - tb, (int64_t)a.L) : 0;
    m.pa[it] = ok ? tab.a0[gg] + tb * at + i : 0;
  }
#pragma unroll
  for (int it = 0; it < SM::ITX; ++it) {
    const int idx = it * NT + (int)threadIdx.x;
    const int gg = min(idx, NX - 1) / C::TB;
    const int tb = min(idx, NX - 1) - gg * C::TB;
    const int64_t rem = tab.rem[gg];
    const bool ok = idx < NX && rem > tb;
    m.limx[it] = ok ? (int)min(rem - tb, (int64_t)a.L) : 0;
    m.px[it] = ok ? tab.x0[gg] + tb : 0;
  }
}

// Staging helper: copy the A rows and the per-step input of the G items for the
// time block [t0, t0 + TB) into LDS; rows beyond the signal end and lags >= n
// are zero. Every thread first issues all of its loads, unconditionally (an
// element that must read as zero loads a valid address and is selected away),
// then writes LDS: a load under a branch is waited for at once, which had
// serialised the thread's loads. The kernels run C::THREADS threads.
template <typename S, int NP, int LANES>
__device__ __forceinline__ void stage_block(const ChunkArgs<S>& a, typename Wide<S>::type* sA,
                                            typename Wide<S>::type* sX, const S* inp,
                                            const StageMap<S, NP, LANES>& m, int t0) {
  using C = ChunkCfg<S, NP, LANES>;
  using W = typename Wide<S>::type;
  constexpr int NT = C::THREADS;
  constexpr int NA = C::G * C::TB * NP;
  constexpr int NX = C::G * C::TB;
  constexpr int ITA = (NA + NT - 1) / NT;
  constexpr int ITX = (NX + NT - 1) / NT;
  S va[ITA], vx[ITX];
  bool oka[ITA], okx[ITX];
  const int64_t toff = (int64_t)t0 * (a.rg ? a.n : a.a_tstride);
#pragma unroll
  for (int it = 0; it < ITA; ++it) {
    const bool ok = t0 < m.lima[it];
    va[it] = a.A[ok ? m.pa[it] + toff : 0];
    oka[it] = ok;
  }
#pragma unroll
  for (int it = 0; it < ITX; ++it) {
    const bool ok = t0 < m.limx[it];
    vx[it] = inp[ok ? m.px[it] + t0 : 0];
    okx[it] = ok;
  }
#pragma unroll
  for (int it = 0; it < ITA; ++it) {
    const int idx = it * NT + (int)threadIdx.x;
    if (idx < NA) {
      const int gg = idx / (C::TB * NP);
      const int r = idx - gg * (C::TB * NP);
      sA[gg * C::ASTRIDE + r] = oka[it] ? convert<W>(coef(a, va[it])) : Num<W>::zero();
    }
  }
#pragma unroll
  for (int it = 0; it < ITX; ++it) {
    const int idx = it * NT + (int)threadIdx.x;
    if (idx < NX) sX[idx] = okx[it] ? convert<W>(vx[it]) : Num<W>::zero();
  }
}

// ----------------------------------------------------------------------------
// Phase 1 (forward): summaries of chunks c = 0 .. P-2 of every signal, NP lanes
// per chunk. Lane k sweeps the chunk backwards with the transposed recurrence
// from a unit at output k (the sample L-1-k of the chunk): the pending sums left
// for the positions before the chunk are d y_{L-1-k} / d y_{-1-j} = Phi[k][j]
// (complex-linear: no conjugate), and the lambdas met on the way are
// d y_{L-1-k} / d x_t, so z[k] = sum_t lambda_t x_t. Each step's NP updates are
// independent of each other (the forward form chains NP products per step), and
// no extra lane carries the zero-state response.
template <typename S, int NP>
__global__ __launch_bounds__(256) void fwd_summary(ChunkArgs<S> a) {
  using C = ChunkCfg<S, NP, NP>;
  using W = typename Wide<S>::type;
  __shared__ W sA[C::G * C::ASTRIDE];
  __shared__ W sX[C::G * C::TB];
  __shared__ ItemTab<S, NP, NP> tab;
  const int64_t first = (int64_t)blockIdx.x * C::G;
  fill_items<S, NP, NP>(a, tab, first, true);
  __syncthreads();
  StageMap<S, NP, NP> sm;
  stage_map<S, NP, NP>(a, sm, tab);
  const int g = threadIdx.x / NP;
  const int k = threadIdx.x - g * NP;
  const bool active = g < C::G && tab.rem[g] > 0;

  W ring[NP];  // pending sums, slot u mod NP for position u
#pragma unroll
  for (int r = 0; r < NP; ++r) ring[r] = (r == NP - 1 - k) ? Num<W>::one() : Num<W>::zero();
  W z = Num<W>::zero();

  for (int t0 = a.L - C::TB; t0 >= 0; t0 -= C::TB) {
    stage_block<S, NP, NP>(a, sA, sX, a.x, sm, t0);
    __syncthreads();
    if (active) {
      const W* rowbase = sA + g * C::ASTRIDE;
#pragma unroll
      for (int jj = 0; jj < C::TB; ++jj) {
        const int j = C::TB - 1 - jj;  // local time; slot of position t is j mod NP
        const W* row = rowbase + j * NP;
        const int h = j & (NP - 1);
        const W lam = ring[h];
        ring[h] = Num<W>::zero();
#pragma unroll
        for (int i = 1; i <= NP; ++i)
          ring[(j - i) & (NP - 1)] = Num<W>::fms(ring[(j - i) & (NP - 1)], row[i - 1], lam);
        z = Num<W>::fma(z, lam, sX[g * C::TB + j]);
      }
    }
    __syncthreads();
  }
  if (active) {
    W* out = a.summ + (first + g) * C::SUMM;
    // pending for position -1-r (before the chunk) sits in slot NP-1-r
#pragma unroll
    for (int r = 0; r < NP; ++r) out[k * NP + r] = ring[NP - 1 - r];
    out[NP * NP + k] = z;
  }
}

// Phase 1 (adjoint): summaries of chunks c = 1 .. P-1 (reverse time).
template <typename S, int NP>
__global__ __launch_bounds__(256) void bwd_summary(ChunkArgs<S> a) {
  using C = ChunkCfg<S, NP>;
  using W = typename Wide<S>::type;
  __shared__ W sA[C::G * C::ASTRIDE];
  __shared__ W sX[C::G * C::TB];
  __shared__ ItemTab<S, NP> tab;
  const int64_t first = (int64_t)blockIdx.x * C::G;
  fill_items<S, NP, NP + 1>(a, tab, first, false);
  __syncthreads();
  StageMap<S, NP, NP + 1> sm;
  stage_map<S, NP, NP + 1>(a, sm, tab);
  const int g = threadIdx.x / (NP + 1);
  const int k = threadIdx.x - g * (NP + 1);
  const bool active = g < C::G && tab.rem[g] > 0;

  W ring[NP];
#pragma unroll
  for (int r = 0; r < NP; ++r) ring[r] = Num<W>::zero();
  // p_in = e_k: pending sum for position t1-1-k, slot (t1-1-k) mod NP = NP-1-k.
#pragma unroll
  for (int r = 0; r < NP; ++r)
    if (k == r) ring[NP - 1 - r] = Num<W>::one();
  const W gmask = (k == NP) ? Num<W>::one() : Num<W>::zero();

  for (int t0 = a.L - C::TB; t0 >= 0; t0 -= C::TB) {
    stage_block<S, NP, NP + 1>(a, sA, sX, a.x, sm, t0);
    __syncthreads();
    if (active) {
      const W* rowbase = sA + g * C::ASTRIDE;
#pragma unroll
      for (int jj = 0; jj < C::TB; ++jj) {
        const int j = C::TB - 1 - jj;  // local time; slot of position t is j mod NP
        const W* row = rowbase + j * NP;
        const int h = j & (NP - 1);
        const W lam = sX[g * C::TB + j] * gmask + ring[h];
        ring[h] = Num<W>::zero();
#pragma unroll
        for (int i = 1; i <= NP; ++i)
          ring[(j - i) & (NP - 1)] = Num<W>::fms(ring[(j - i) & (NP - 1)], Num<W>::conj(row[i - 1]), lam);
      }
    }
    __syncthreads();
  }
  if (active) {
    W* out = a.summ + (first + g) * C::SUMM;
    // final pending for position t0-1-r sits in slot NP-1-r
    if (k < NP) {
#pragma unroll
      for (int r = 0; r < NP; ++r) out[r * NP + k] = ring[NP - 1 - r];
    } else {
#pragma unroll
      for (int r = 0; r < NP; ++r) out[NP * NP + r] = ring[NP - 1 - r];
    }
  }
}

// Phase 1 (adjoint): summaries of chunks c = 1 .. P-1, NP lanes per chunk, in
// forward time. Lane r runs the forward recurrence with zero input from a unit
// at position t0-1-r (the state entering the chunk): its outputs h_r(t) are
// d y_t / d y_{t0-1-r}. The adjoint chunk map is the conjugate transpose of the
// forward one, so its matrix row r is Psi[r][k] = conj(h_r(t1-1-k)) (the lane's
// final window), and its zero-state response to g is w[r] = sum_t g_t conj(h_r(t)).
// One pass gives both. The adjoint-form kernel above needs an extra lane per
// chunk for w, and at order 32 its 33-lane groups straddle the waves (measured
// in the ragged batch: 1248 -> 606 us); at the lower orders it is the faster
// of the two, so this form serves order 32 only. Positions past the
// end of the signal have zero rows and g = 0 (staged as zeros), so they add
// nothing to w; Psi of a signal's last chunk meets the zero state.
template <typename S, int NP>
__global__ __launch_bounds__(256) void bwd_summary_ff(ChunkArgs<S> a) {
  using C = ChunkCfg<S, NP, NP>;
  using W = typename Wide<S>::type;
  __shared__ W sA[C::G * C::ASTRIDE];
  __shared__ W sX[C::G * C::TB];
  __shared__ ItemTab<S, NP, NP> tab;
  const int64_t first = (int64_t)blockIdx.x * C::G;
  fill_items<S, NP, NP>(a, tab, first, false);
  __syncthreads();
  StageMap<S, NP, NP> sm;
  stage_map<S, NP, NP>(a, sm, tab);
  const int g = threadIdx.x / NP;
  const int r = threadIdx.x - g * NP;
  const bool active = g < C::G && tab.rem[g] > 0;

  W ring[NP];  // h_r(u) in slot u mod NP
#pragma unroll
  for (int k = 0; k < NP; ++k) ring[k] = (k == NP - 1 - r) ? Num<W>::one() : Num<W>::zero();
  W w = Num<W>::zero();

  for (int t0 = 0; t0 < a.L; t0 += C::TB) {
    stage_block<S, NP, NP>(a, sA, sX, a.x, sm, t0);
    __syncthreads();
    if (active) {
      const W* rowbase = sA + g * C::ASTRIDE;
#pragma unroll
      for (int j = 0; j < C::TB; ++j) {
        const W* row = rowbase + j * NP;
        // four partial sums keep the step's dependency chain at NP/4 products
        W acc0 = Num<W>::zero(), acc1 = Num<W>::zero(), acc2 = Num<W>::zero(), acc3 = Num<W>::zero();
#pragma unroll
        for (int i = 1; i <= NP; ++i) {
          W& acc = (i & 3) == 1 ? acc0 : (i & 3) == 2 ? acc1 : (i & 3) == 3 ? acc2 : acc3;
          acc = Num<W>::fms(acc, row[i - 1], ring[(j - i) & (NP - 1)]);
        }
        const W y = (acc0 + acc1) + (acc2 + acc3);
        ring[j & (NP - 1)] = y;
        w = Num<W>::fma(w, sX[g * C::TB + j], Num<W>::conj(y));
      }
    }
    __syncthreads();
  }
  if (active) {
    W* out = a.summ + (first + g) * C::SUMM;
#pragma unroll
    for (int k = 0; k < NP; ++k) out[r * NP + k] = Num<W>::conj(ring[NP - 1 - k]);
    out[NP * NP + r] = w;
  }
}

// ----------------------------------------------------------------------------
// Phase 2: carries, per signal.
// Forward: carry[c] = state entering chunk c, carry[0] = zi, carry[c+1] = Phi_c carry[c] + z_c.
// Adjoint: carry[c] = pending sums entering chunk c from the right,
//          carry[P-1] = 0, carry[c-1] = Psi_c carry[c] + w_c (summary of chunk c at c-1).
// Pipelined carry: one 256-thread workgroup per signal. Wave 0 walks the chain;
// waves 1-3 take turns staging batches of CB links from the summaries into
// registers DEPTH batches ahead and committing them into an LDS double buffer
// (matrix rows padded to NP+1 elements), one LDS barrier per batch. A link is
// then LDS-bound only: the chain lane r reads row r of M and the broadcast state,
// and the next link's row is read before the current link's products.
template <typename S, int NP>
struct PipeCfg {
  using W = typename Wide<S>::type;
  static constexpr int ROWP = NP + 1;
  static constexpr int LINK = NP * ROWP + NP;  // staged elements per link
  static constexpr int SUMM = NP * NP + NP;    // summary elements per link
  static constexpr int BUDGET = 72 * 1024;
  static constexpr int CB = (2 * 8 * LINK * (int)sizeof(W) <= BUDGET) ? 8
                            : ((2 * 4 * LINK * (int)sizeof(W) <= BUDGET) ? 4 : 2);
  static constexpr int PER = (CB * SUMM + 63) / 64;  // elements per loader lane per batch
  static constexpr int DEPTH = 3;
};

template <typename S, int NP, bool FWD>
__global__ __launch_bounds__(256) void carry_pipe(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  using C = PipeCfg<S, NP>;
  __shared__ W buf[2][C::CB][C::LINK];
  __shared__ W sv[NP];
  const int64_t b = blockIdx.x;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const Span sp = sig_span(a, b);
  const int64_t P = sp.P;
  const int64_t links = P - 1;
  const int64_t nbatch = (links + C::CB - 1) / C::CB;
  W* carry = a.carry + sp.cbase * NP;
  W stage[C::PER];
  // link l (in processing order) is chunk FWD ? l : P-1-l
  // loads are unconditional (indices clamped to the last link: the chain never
  // reads those slots), so that nothing waits for them before the commit
  auto fetch = [&](int64_t j) {
    const int64_t l0 = j * C::CB;
#pragma unroll
    for (int q = 0; q < C::PER; ++q) {
      const int e = min(q * 64 + lane, C::CB * C::SUMM - 1);
      const int li = e / C::SUMM, off = e - li * C::SUMM;
      const int64_t l = min(l0 + li, links - 1);
      const int64_t c = FWD ? l : P - 1 - l;
      stage[q] = a.summ[summ_slot(a, b, sp, c, FWD) * C::SUMM + off];
    }
  };
  // An adjoint handed the forward's summaries (a.kept) reads each chunk's Phi
  // as Psi = Phi^T, transposing while it commits, and the matrix of its first
  // link (the signal's last chunk, which has no forward summary) as zero: that
  // link meets the zero state, and the slot holds whatever was there.
  const bool tr = !FWD && a.kept;
  auto commit = [&](int64_t j) {
    W(*dst)[C::LINK] = buf[j & 1];
#pragma unroll
    for (int q = 0; q < C::PER; ++q) {
      const int e = q * 64 + lane;
      const int li = e / C::SUMM, off = e - li * C::SUMM;
      if (e < C::CB * C::SUMM) {
        if (off >= NP * NP) {
          dst[li][off + NP] = stage[q];
        } else if (!tr) {
          dst[li][off + off / NP] = stage[q];
        } else {
          const int k = off / NP, r = off - k * NP;  // Phi[k][r] is Psi[r][k]
          dst[li][r * C::ROWP + k] = (j * C::CB + li == 0) ? Num<W>::zero() : stage[q];
        }
      }
    }
  };
  const bool own = wave == 0 && lane < NP;
  // float64 chains of order 8, 16 and 32 take the link's products through DPP
  // row broadcasts of the state (no LDS round trip per link): orders 8 and 16
  // on 16-lane rows (lane r owns s_r and row r; at order 8 lanes 8-15 repeat
  // lanes 0-7), order 32 on 32-lane halves (lane e = 16h + l owns s_e and columns
  // 16h..16h+15 of rows l and l + 16); the other lanes of the wave repeat them.
  // Other orders and types broadcast the state through LDS.
  constexpr bool kDpp = std::is_same<W, double>::value && (NP == 8 || NP == 16 || NP == 32);
  const int de = NP == 32 ? (lane & 31) : (lane & (NP - 1));  // DPP form: the state element of the lane
  W s = Num<W>::zero();
  if constexpr (kDpp) {
    if (FWD && wave == 0 && de < a.n && a.zi) s = convert<W>(a.zi[sp.zo + de]);
  } else {
    if (FWD && own && lane < a.n && a.zi) s = convert<W>(a.zi[sp.zo + lane]);
  }
  const int row = lane < NP ? lane : NP - 1;
  for (int64_t i = -C::DEPTH; i < nbatch; ++i) {
    if (wave > 0) {
      const int64_t jn = i + C::DEPTH;  // issue
      if (jn < nbatch && 1 + jn % 3 == wave) fetch(jn);
      const int64_t jc = i + 1;  // commit (fetched DEPTH - 1 iterations ago)
      if (jc >= 0 && jc < nbatch && 1 + jc % 3 == wave) commit(jc);
    } else if (i >= 0 && kDpp) {
     if constexpr (kDpp) {
      const int64_t l0 = i * C::CB;
      const int nb = (int)((links - l0) < C::CB ? (links - l0) : C::CB);
      const W(*src)[C::LINK] = buf[i & 1];
      if constexpr (NP == 32) {
        const int h = (lane >> 4) & 1, l = lane & 15;
        double ma[2][16], mb[2][16], zv[2];
        auto ldm = [&](int li, double(&a0)[16], double(&b0)[16], double& z) {
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            a0[j] = src[li][l * C::ROWP + 16 * h + j];
            b0[j] = src[li][(l + 16) * C::ROWP + 16 * h + j];
          }
          z = src[li][NP * C::ROWP + de];
        };
        ldm(0, ma[0], mb[0], zv[0]);
#pragma unroll
        for (int li = 0; li < C::CB; ++li) {
          if (li + 1 < C::CB) ldm(li + 1, ma[(li + 1) & 1], mb[(li + 1) & 1], zv[(li + 1) & 1]);
          if (li < nb) {
            const int64_t l1 = l0 + li;
            if (own) carry[(FWD ? l1 : P - 1 - l1) * NP + lane] = s;
            s = dpp16::dpp_matvec32(s, ma[li & 1], mb[li & 1], h) + zv[li & 1];
          }
        }
      } else {
        double m[2][NP], zv[2];
        auto ldm = [&](int li, double(&mm)[NP], double& z) {
#pragma unroll
          for (int k = 0; k < NP; ++k) mm[k] = src[li][de * C::ROWP + k];
          z = src[li][NP * C::ROWP + de];
        };
        ldm(0, m[0], zv[0]);
#pragma unroll
        for (int li = 0; li < C::CB; ++li) {
          if (li + 1 < C::CB) ldm(li + 1, m[(li + 1) & 1], zv[(li + 1) & 1]);
          if (li < nb) {
            const int64_t l1 = l0 + li;
            if (own) carry[(FWD ? l1 : P - 1 - l1) * NP + lane] = s;
            if constexpr (NP == 16)
              s = dpp16::chain_link(s, m[li & 1], zv[li & 1]);
            else
              s = dpp16::chain_link8(s, m[li & 1], zv[li & 1]);
          }
        }
      }
     }
    } else if (i >= 0) {
      const int64_t l0 = i * C::CB;
      const int nb = (int)((links - l0) < C::CB ? (links - l0) : C::CB);
      const W(*src)[C::LINK] = buf[i & 1];
      W m[2][NP];
      W zv[2];
      auto ldm = [&](int li, W(&mm)[NP], W& z) {
#pragma unroll
        for (int k = 0; k < NP; ++k) mm[k] = src[li][row * C::ROWP + k];
        z = src[li][NP * C::ROWP + row];
      };
      ldm(0, m[0], zv[0]);
#pragma unroll
      for (int li = 0; li < C::CB; ++li) {
        if (li + 1 < C::CB) ldm(li + 1, m[(li + 1) & 1], zv[(li + 1) & 1]);
        if (li < nb) {
          const int64_t l = l0 + li;
          if (own) {
            carry[(FWD ? l : P - 1 - l) * NP + lane] = s;
            sv[lane] = s;
          }
          // the state reaches every lane through LDS (one wave: in order)
          W acc0 = zv[li & 1], acc1 = Num<W>::zero(), acc2 = Num<W>::zero(), acc3 = Num<W>::zero();
#pragma unroll
          for (int k = 0; k < NP; k += 4) {
            acc0 = Num<W>::fma(acc0, m[li & 1][k], sv[k]);
            if (k + 1 < NP) acc1 = Num<W>::fma(acc1, m[li & 1][k + 1], sv[k + 1]);
            if (k + 2 < NP) acc2 = Num<W>::fma(acc2, m[li & 1][k + 2], sv[k + 2]);
            if (k + 3 < NP) acc3 = Num<W>::fma(acc3, m[li & 1][k + 3], sv[k + 3]);
          }
          s = (acc0 + acc1) + (acc2 + acc3);
        }
      }
    }
    // LDS-only barrier: the loaders' global loads stay in flight across it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
  if (own) carry[(FWD ? P - 1 : 0) * NP + lane] = s;
}

// Blocked carry scan for small NP: one workgroup per signal; each thread composes
// the affine maps of a contiguous run of chunks, the runs are combined by a
// workgroup scan, then each thread re-walks its run from its carry.
template <typename S, int NP>
struct Affine {
  S M[NP][NP];
  S v[NP];
};

template <typename S, int NP>
__device__ __forceinline__ void affine_identity(Affine<S, NP>& f) {
#pragma unroll
  for (int r = 0; r < NP; ++r) {
#pragma unroll
    for (int k = 0; k < NP; ++k) f.M[r][k] = (r == k) ? Num<S>::one() : Num<S>::zero();
    f.v[r] = Num<S>::zero();
  }
}
// out = second o first  (apply first, then second)
template <typename S, int NP>
__device__ __forceinline__ Affine<S, NP> affine_compose(const Affine<S, NP>& second,
                                                        const Affine<S, NP>& first) {
  Affine<S, NP> o;
#pragma unroll
  for (int r = 0; r < NP; ++r) {
    S vv = second.v[r];
#pragma unroll
    for (int m = 0; m < NP; ++m) vv = Num<S>::fma(vv, second.M[r][m], first.v[m]);
    o.v[r] = vv;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      S acc = Num<S>::zero();
#pragma unroll
      for (int m = 0; m < NP; ++m) acc = Num<S>::fma(acc, second.M[r][m], first.M[m][k]);
      o.M[r][k] = acc;
    }
  }
  return o;
}

template <typename S, int NP>
__device__ __forceinline__ void load_affine(Affine<S, NP>& f, const S* sm) {
#pragma unroll
  for (int r = 0; r < NP; ++r) {
#pragma unroll
    for (int k = 0; k < NP; ++k) f.M[r][k] = sm[r * NP + k];
    f.v[r] = sm[NP * NP + r];
  }
}

constexpr int kBlockedThreads = 256;

// Long chains of a uniform batch are split into a.nseg segments of links, one
// workgroup each: the per-link loads of one workgroup are scattered (a thread's
// run of links is contiguous, so a wave's loads touch 64 lines each), and a
// signal's whole chain in one workgroup was bound by that one CU's memory
// pipeline. SEG_REDUCE composes a segment's links into one map, carry_segs
// chains the maps of each signal, and SEG_APPLY re-walks every segment from its
// incoming state. SEG_WHOLE walks a signal's chain in one workgroup.
// Workspace a.segs: the maps [B][nseg][NP*NP+NP], then the incoming states
// [B][nseg][NP].
enum { SEG_WHOLE = 0, SEG_REDUCE = 1, SEG_APPLY = 2 };

template <typename S, int NP, bool FWD, int MODE>
__global__ __launch_bounds__(kBlockedThreads) void carry_blocked(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  __shared__ Affine<W, NP> sh[kBlockedThreads];
  const int64_t b = MODE == SEG_WHOLE ? blockIdx.x : blockIdx.y;
  const int tid = threadIdx.x;
  const Span sp = sig_span(a, b);
  const int64_t P = sp.P;
  const int64_t links = P - 1;  // number of affine maps in the chain
  int64_t l0 = 0, l1 = links;   // this workgroup's links
  if constexpr (MODE != SEG_WHOLE) {
    const int64_t seg = (links + a.nseg - 1) / a.nseg;
    l0 = min((int64_t)blockIdx.x * seg, links);
    l1 = min(l0 + seg, links);
  }
  const int64_t per = (l1 - l0 + kBlockedThreads - 1) / kBlockedThreads;
  const int64_t lo = min(l0 + tid * per, l1);
  const int64_t hi = min(lo + per, l1);
  // link l (0-based, in processing order) maps carry of step l to carry of step l+1;
  // forward: chunk c = l; adjoint: chunk c = P-1-l.
  auto link = [&](int64_t l) -> const W* {
    return a.summ + summ_slot(a, b, sp, FWD ? l : P - 1 - l, FWD) * ChunkCfg<S, NP>::SUMM;
  };
  // A thread walks its run in batches of UR links whose loads are issued
  // together (at clamped link indices, so unconditionally) before the batch's
  // arithmetic: the chain then waits for memory once per batch, not per link.
  constexpr int UR = NP <= 2 ? 8 : 4;
  Affine<W, NP> f;
  affine_identity(f);
  for (int64_t l = lo; l < hi; l += UR) {
    Affine<W, NP> m[UR];
#pragma unroll
    for (int u = 0; u < UR; ++u) load_affine(m[u], link(l + u < hi ? l + u : hi - 1));
#pragma unroll
    for (int u = 0; u < UR; ++u)
      if (l + u < hi) f = affine_compose(m[u], f);
  }
  sh[tid] = f;
  __syncthreads();
  // inclusive Hillis-Steele scan over threads (composition in thread order)
  for (int off = 1; off < kBlockedThreads; off <<= 1) {
    Affine<W, NP> other;
    const bool has = tid >= off;
    if (has) other = sh[tid - off];
    __syncthreads();
    if (has) {
      f = affine_compose(f, other);
      sh[tid] = f;
    }
    __syncthreads();
  }
  if constexpr (MODE == SEG_REDUCE) {
    // the last thread's inclusive composition is the segment's map
    if (tid == kBlockedThreads - 1) {
      W* out = a.segs + (b * a.nseg + blockIdx.x) * (NP * NP + NP);
#pragma unroll
      for (int r = 0; r < NP; ++r) {
#pragma unroll
        for (int k = 0; k < NP; ++k) out[r * NP + k] = f.M[r][k];
        out[NP * NP + r] = f.v[r];
      }
    }
    return;
  }
  // exclusive prefix applied to the initial carry (the segment's incoming state)
  W s[NP];
  if constexpr (MODE == SEG_APPLY) {
    const W* in = a.segs + a.B * a.nseg * (NP * NP + NP) + (b * a.nseg + blockIdx.x) * NP;
#pragma unroll
    for (int r = 0; r < NP; ++r) s[r] = in[r];
  } else {
#pragma unroll
    for (int r = 0; r < NP; ++r)
      s[r] = (FWD && r < a.n && a.zi) ? convert<W>(a.zi[sp.zo + r]) : Num<W>::zero();
  }
  if (tid > 0) {
    const Affine<W, NP>& pre = sh[tid - 1];
    W t[NP];
#pragma unroll
    for (int r = 0; r < NP; ++r) {
      W acc = pre.v[r];
#pragma unroll
      for (int k = 0; k < NP; ++k) acc = Num<W>::fma(acc, pre.M[r][k], s[k]);
      t[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < NP; ++r) s[r] = t[r];
  }
  W* carry = a.carry + sp.cbase * NP;
  for (int64_t l = lo; l < hi; l += UR) {
    Affine<W, NP> m[UR];
#pragma unroll
    for (int u = 0; u < UR; ++u) load_affine(m[u], link(l + u < hi ? l + u : hi - 1));
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      if (l + u < hi) {
        const int64_t c = FWD ? l + u : P - 1 - (l + u);
#pragma unroll
        for (int r = 0; r < NP; ++r) carry[c * NP + r] = s[r];
        W t[NP];
#pragma unroll
        for (int r = 0; r < NP; ++r) {
          W acc = m[u].v[r];
#pragma unroll
          for (int k = 0; k < NP; ++k) acc = Num<W>::fma(acc, m[u].M[r][k], s[k]);
          t[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < NP; ++r) s[r] = t[r];
      }
    }
  }
  // the thread owning the last link (thread 0 when there is none) writes the final carry
  if ((lo < hi && hi == links) || (links == 0 && tid == 0)) {
    const int64_t c = FWD ? P - 1 : 0;
#pragma unroll
    for (int r = 0; r < NP; ++r) carry[c * NP + r] = s[r];
  }
}

// Segmented carry, middle step: per signal, the state entering each segment
// (one workgroup per signal; thread s loads the map of segment s, then thread 0
// walks the maps in LDS). At most kBlockedThreads segments.
template <typename S, int NP, bool FWD>
__global__ __launch_bounds__(kBlockedThreads) void carry_segs(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  __shared__ Affine<W, NP> sh[kBlockedThreads];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const W* maps = a.segs + b * a.nseg * (NP * NP + NP);
  if (tid < a.nseg) load_affine(sh[tid], maps + tid * (NP * NP + NP));
  __syncthreads();
  if (tid != 0) return;
  const Span sp = sig_span(a, b);
  W s[NP];
#pragma unroll
  for (int r = 0; r < NP; ++r)
    s[r] = (FWD && r < a.n && a.zi) ? convert<W>(a.zi[sp.zo + r]) : Num<W>::zero();
  W* in = a.segs + a.B * a.nseg * (NP * NP + NP) + b * a.nseg * NP;
  for (int64_t g = 0; g < a.nseg; ++g) {
#pragma unroll
    for (int r = 0; r < NP; ++r) in[g * NP + r] = s[r];
    W t[NP];
#pragma unroll
    for (int r = 0; r < NP; ++r) {
      W acc = sh[g].v[r];
#pragma unroll
      for (int k = 0; k < NP; ++k) acc = Num<W>::fma(acc, sh[g].M[r][k], s[k]);
      t[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < NP; ++r) s[r] = t[r];
  }
}

// Launches the carry chain of every signal with the blocked scan: segmented
// when the plan gave a segment workspace (a.nseg > 0), else one workgroup per
// signal.
template <typename S, int NP, bool FWD>
void launch_carry_blocked(const ChunkArgs<S>& a, hipStream_t st) {
  if (a.nseg > 0) {
    const dim3 g((unsigned)a.nseg, (unsigned)a.B);
    hipLaunchKernelGGL((carry_blocked<S, NP, FWD, SEG_REDUCE>), g, dim3(kBlockedThreads), 0, st, a);
    hipLaunchKernelGGL((carry_segs<S, NP, FWD>), dim3((unsigned)a.B), dim3(kBlockedThreads), 0, st, a);
    hipLaunchKernelGGL((carry_blocked<S, NP, FWD, SEG_APPLY>), g, dim3(kBlockedThreads), 0, st, a);
  } else {
    hipLaunchKernelGGL((carry_blocked<S, NP, FWD, SEG_WHOLE>), dim3((unsigned)a.B), dim3(kBlockedThreads), 0,
                       st, a);
  }
}

// The blocked scan keeps one affine map per thread in LDS; it is used where that
// fits in 48 KiB.
template <typename S, int NP>
constexpr bool blocked_fits() {
  return (size_t)NP * (NP + 1) * sizeof(typename Wide<S>::type) * kBlockedThreads <= 48 * 1024;
}

// ----------------------------------------------------------------------------
// Row loader: the NP coefficients of one time step, zero beyond the order. It
// reads every slot at a valid address (lag min(i, n-1)) and selects
// zero afterwards, so that no load depends on a branch.
// The zeroing is a bit mask, not a select: the code generator turns a select
// of a loaded value into a branch and sinks the load into it.
__device__ __forceinline__ float keep_bits(float v, bool k) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & (k ? 0xffffffffu : 0u));
}
__device__ __forceinline__ double keep_bits(double v, bool k) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) & (k ? ~0ull : 0ull));
}
template <typename R>
__device__ __forceinline__ Cx<R> keep_bits(Cx<R> v, bool k) {
  return {keep_bits(v.re, k), keep_bits(v.im, k)};
}
template <typename S, int NP>
__device__ __forceinline__ void load_row_u(S (&row)[NP], const ChunkArgs<S>& a, const S* p) {
#pragma unroll
  for (int i = 0; i < NP; ++i) row[i] = keep_bits(coef(a, p[i < a.n ? i : a.n - 1]), i < a.n);
}
// The plain form, for steps that sit under a branch of their own. Float32 rows
// of a multiple-of-4 order that start on a 16-byte boundary (always in a
// uniform batch with an aligned A; per signal in a ragged one) load as float4:
// a lane's 4-byte loads are L2 requests of their own.
template <typename S, int NP>
__device__ __forceinline__ void load_row(S (&row)[NP], const ChunkArgs<S>& a, const S* p) {
  if constexpr (std::is_same<S, float>::value && NP % 4 == 0) {
    if (a.n == NP && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
      const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
      for (int k = 0; k < NP / 4; ++k) {
        const float4 v = q[k];
        row[4 * k] = coef(a, v.x);
        row[4 * k + 1] = coef(a, v.y);
        row[4 * k + 2] = coef(a, v.z);
        row[4 * k + 3] = coef(a, v.w);
      }
      return;
    }
  }
  if (a.n == NP) {
#pragma unroll
    for (int i = 0; i < NP; ++i) row[i] = coef(a, p[i]);
  } else {
#pragma unroll
    for (int i = 0; i < NP; ++i) row[i] = (i < a.n) ? coef(a, p[i]) : Num<S>::zero();
  }
}

// ----------------------------------------------------------------------------
// Phase 3 (reruns), time-invariant filters: one lane per chunk re-filters its
// samples from the carry, in the wide type (float64 for float32 data), so the
// only single-precision rounding of an output is its final store.
//
// The coefficients are loaded once. The per-sample streams (x or g, and y for
// da) are prefetched DS steps ahead in rings whose slots are consumed before
// they are reloaded. Every load is unconditional, at an index clamped into the
// signal, and values that must count as zero are selected after the load; a
// load under a branch would leave the compiler a phi that it waits for at once.
// Steps past the end of a signal compute on clamped data and store nothing
// (forward) or are no-ops of a zero lambda (adjoint).
template <typename S, int NP>
struct RerunCfg {
  static constexpr int K = NP < 16 ? 16 : NP;  // steps per unrolled block: a multiple of NP and DS
  static constexpr int DS = 16;
};

__device__ __forceinline__ int64_t clamp64(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Workgroup barrier that orders LDS only: __syncthreads() also fences global
// memory, which waits for every load in flight (a prefetch ring).
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// float4 streams of the time-invariant float32 reruns: component K of a block of
// four consecutive samples, loaded and stored whole.
template <int K>
__device__ __forceinline__ float v4get(const float4& v) {
  if constexpr (K == 0) return v.x;
  if constexpr (K == 1) return v.y;
  if constexpr (K == 2) return v.z;
  return v.w;
}
template <int K>
__device__ __forceinline__ void v4set(float4& v, float f) {
  if constexpr (K == 0) v.x = f;
  if constexpr (K == 1) v.y = f;
  if constexpr (K == 2) v.z = f;
  if constexpr (K == 3) v.w = f;
}
__device__ __forceinline__ float4 ld4(const float* p, int64_t i) {
  return *reinterpret_cast<const float4*>(p + i);
}
constexpr int floordiv4(int v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); }
constexpr int posmod(int v, int m) { return ((v % m) + m) % m; }

// Keeps a step's loads where the source puts them (see tlpc_chunk16.hip).
__device__ __forceinline__ void pin_step() { __builtin_amdgcn_sched_barrier(0); }

// The NP coefficients as the recurrence uses them: -A for the first-order scan,
// zero at and past the order (those slots read a valid element first).
template <typename S, int NP>
__device__ __forceinline__ void load_coefs_w(typename Wide<S>::type (&cw)[NP], const ChunkArgs<S>& a, const S* p) {
  using W = typename Wide<S>::type;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const W v = convert<W>(p[i < a.n ? i : a.n - 1]);
    const W c = a.neg ? -v : v;
    cw[i] = i < a.n ? c : Num<W>::zero();
  }
}

template <typename T>
__device__ __forceinline__ T opaque_v(T v) {
  static_assert(sizeof(T) % 4 == 0, "register-sized values");
  unsigned int w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) asm volatile("" : "+v"(w[k]));
  __builtin_memcpy(&v, w, sizeof(T));
  return v;
}

// VEC (float32 data, T a multiple of 4, x and y 16-byte aligned): x is read and
// y written as one float4 per four steps instead of one 4-byte access per step;
// each 4-byte access of a lane is a request of its own, and those requests, not
// the bytes, bound the scalar form.
template <typename S, int NP, bool VEC>
struct FwdLaneTI {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  W ring[NP];                        // pending sums: slot u mod NP collects -a_i y_{u-i}
  W cw[NP];                          // coefficients
  S xs[VEC ? 1 : RC::DS];            // x_t in slot t mod DS
  float4 xq[VEC ? RC::DS / 4 : 1];   // VEC: x[4q' .. 4q'+3] in slot q' mod DS/4
  float4 yq;                         // VEC: the outputs of the current four steps
  const S* xb;
  S* yb;
  int64_t T, tend;
};

template <typename S, int NP, bool VEC, int JJ>
__device__ __forceinline__ void fwd_steps_ti(FwdLaneTI<S, NP, VEC>& st, int64_t tb) {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  if constexpr (JJ < RC::K) {
    constexpr int j = JJ % NP;  // t mod NP: blocks start at multiples of NP
    const int64_t t = tb + JJ;
    // pending-sum form: y_t = x_t + p(t), then p(t+i) -= a_i y_t; the step's NP
    // products are independent and only the first feeds the next step (the
    // direct form chained NP products per step)
    W acc;
    if constexpr (VEC)
      acc = convert<W>(v4get<JJ & 3>(st.xq[(JJ % RC::DS) / 4])) + st.ring[j];
    else
      acc = convert<W>(st.xs[JJ % RC::DS]) + st.ring[j];
    st.ring[j] = Num<W>::zero();
#pragma unroll
    for (int i = 1; i <= NP; ++i)
      st.ring[(j + i) & (NP - 1)] = Num<W>::fms(st.ring[(j + i) & (NP - 1)], st.cw[i - 1], acc);
    if constexpr (VEC) {
      // the block's last sample was just read: its slot takes the block DS ahead
      if constexpr ((JJ & 3) == 3) st.xq[(JJ % RC::DS) / 4] = ld4(st.xb, min(t - 3 + RC::DS, st.T - 4));
    } else {
      st.xs[JJ % RC::DS] = st.xb[t + RC::DS < st.T ? t + RC::DS : st.T - 1];
    }
    pin_step();
    if constexpr (VEC) {
      v4set<JJ & 3>(st.yq, (float)acc);
      // chunk ends are multiples of 4, so a block is wholly inside or past the end
      if constexpr ((JJ & 3) == 3)
        if (t < st.tend) *reinterpret_cast<float4*>(st.yb + t - 3) = st.yq;
    } else {
      if (t < st.tend) st.yb[t] = convert<S>(acc);
    }
    fwd_steps_ti<S, NP, VEC, JJ + 1>(st, tb);
  }
}

template <typename S, int NP, bool VEC>
__global__ __launch_bounds__(256) void fwd_rerun_ti(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cid >= a.B * a.P) return;
  const int64_t b = cid / a.P;
  const int64_t c = cid - b * a.P;
  const int64_t t0 = c * a.L;
  const int64_t rem = a.T - t0;
  const int len = rem < a.L ? (int)rem : a.L;
  FwdLaneTI<S, NP, VEC> st;
  st.xb = a.x + b * a.T;
  st.yb = a.y + b * a.T;
  st.T = a.T;
  st.tend = t0 + len;
  if constexpr (VEC) {
#pragma unroll
    for (int q = 0; q < RC::DS / 4; ++q) st.xq[q] = opaque_v(ld4(st.xb, min(t0 + 4 * q, a.T - 4)));
  } else {
#pragma unroll
    for (int k = 0; k < RC::DS; ++k) st.xs[k] = opaque_v(st.xb[t0 + k < a.T ? t0 + k : a.T - 1]);
  }
  load_coefs_w<S, NP>(st.cw, a, a.A + b * a.a_bstride);
  const W* cr = a.carry + cid * NP;
  // the carry is the window y_{t0-1-r} (r < NP); it leaves the pending sums
  // p(t0+u) = -sum_{i > u} a_i y_{t0+u-i} for the chunk's first NP positions
  W win[NP];
#pragma unroll
  for (int r = 0; r < NP; ++r) win[r] = opaque_v(cr[r]);
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    W pu = Num<W>::zero();
#pragma unroll
    for (int i = u + 1; i <= NP; ++i) pu = Num<W>::fms(pu, st.cw[i - 1], win[i - u - 1]);
    st.ring[u] = pu;
  }
  for (int jb = 0; jb < len; jb += RC::K) fwd_steps_ti<S, NP, VEC, 0>(st, t0 + jb);
}

// y_u of one signal for the adjoint's window: zi for u < 0 (k < n), zero past T,
// read at an always-valid address and selected afterwards. Selects, not
// branches: a load whose address comes out of a branch joins the loop state
// through a phi, and the loop latch then copies it and waits for it.
template <typename S>
__device__ __forceinline__ const S* yw_addr(const S* yb, const S* zib, int n, int64_t T, int64_t u) {
  const int64_t uc = u < T ? u : T - 1;
  const int64_t k = -u - 1;
  const S* zp = zib + (k < n ? k : n - 1);
  const S* zq = zib ? zp : yb;
  return u >= 0 ? yb + uc : zq;
}
template <typename S>
__device__ __forceinline__ bool yw_ok(const S* zib, int n, int64_t T, int64_t u) {
  return u >= 0 ? u < T : (zib != nullptr && -u - 1 < n);
}

// One lane's adjoint state and streams (see bwd_rerun_ti). DA_PARTIAL also
// accumulates the lane's share of da.
enum DaMode { DA_NONE = 0, DA_PARTIAL = 2 };

// VEC (float32 data, T a multiple of 4, g and dx 16-byte aligned): g is read and
// dx written as one float4 per four steps.
template <typename S, int NP, bool VEC>
struct BwdLaneTI {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  W ring[NP];    // pending sums, slot u mod NP for position u
  W cw[NP];      // conj of the coefficients
  S yr[NP];      // conj(y) window: slot u mod NP holds y_u for u in [t-NP, t-1] (exact)
  W dacc[NP];    // da partials
  S gs[VEC ? 1 : RC::DS];           // g_t in slot (tstart - t) mod DS
  float4 gq[VEC ? RC::DS / 4 : 1];  // VEC: g[u .. u+3] (u a multiple of 4), by step block
  float4 dq;                        // VEC: lambda of the current four steps
  S ys[VEC ? 1 : RC::DS];           // y_{t-1-NP} (window input of step t), same slots
  float4 yq[VEC ? RC::DS / 4 : 1];  // VEC: y blocks of the window input (positions t-1-NP)
  const float* ypadb;               // VEC: this signal's y values below position 0
  const S* gb;
  const S* yb;
  const S* zib;
  S* dxb;
  int64_t T;
};

// y[u .. u+3] (u a multiple of 4) for the VEC adjoint's window: below position 0
// from the signal's ypad blocks (zi), above from y clamped into the signal
// (values past T meet only steps past the end, whose lambda is zero). The choice
// is made on the address, so the load is never under a branch; blocks below the
// ypad range are never read by a step (they are the prefetch past the walk).
template <typename S, int NP, bool VEC>
__device__ __forceinline__ float4 ywin_ti(const BwdLaneTI<S, NP, VEC>& st, int64_t u) {
  const float* p = u < 0 ? st.ypadb + (u + ti_ypad(NP) > 0 ? u + ti_ypad(NP) : 0)
                         : st.yb + (u < st.T - 4 ? u : st.T - 4);
  return *reinterpret_cast<const float4*>(p);
}

// Steps t = tb - JJ for JJ = 0..K-1 (t mod NP = NP-1-JJ mod NP), unrolled by
// template recursion so that every ring index is a compile-time constant (a
// runtime index would put the float64 ring in scratch memory). DA is a template
// argument: a runtime choice inside the steps splits them into branches, and the
// wait-count pass then drains the loads in flight.
template <typename S, int NP, int DA, bool VEC, int JJ>
__device__ __forceinline__ void bwd_steps_ti(const ChunkArgs<S>& a, BwdLaneTI<S, NP, VEC>& st, int64_t tb) {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  if constexpr (JJ < RC::K) {
    constexpr int j = (NP - 1 - JJ % NP);
    const int64_t t = tb - JJ;  // tb is 3 mod 4: step JJ reads component 3 - JJ % 4
    const bool live = t < st.T;
    W gval;
    if constexpr (VEC)
      gval = convert<W>(v4get<3 - (JJ & 3)>(st.gq[(JJ % RC::DS) / 4]));
    else
      gval = convert<W>(st.gs[JJ % RC::DS]);
    const W lam0 = gval + st.ring[j];
    const W lam = live ? lam0 : Num<W>::zero();
    st.ring[j] = live ? Num<W>::zero() : st.ring[j];
#pragma unroll
    for (int i = 1; i <= NP; ++i)
      st.ring[(j - i) & (NP - 1)] = Num<W>::fms(st.ring[(j - i) & (NP - 1)], st.cw[i - 1], lam);
    // the slot just consumed takes the step (VEC: the block) DS further down
    if constexpr (VEC) {
      if constexpr ((JJ & 3) == 3) st.gq[(JJ % RC::DS) / 4] = ld4(st.gb, clamp64(t - RC::DS, 0, st.T - 4));
      v4set<3 - (JJ & 3)>(st.dq, (float)lam);
      // t is a multiple of 4 here and T is too: the block is all live or all past T
      if constexpr ((JJ & 3) == 3)
        if (live && st.dxb) *reinterpret_cast<float4*>(st.dxb + t) = st.dq;
    } else {
      st.gs[JJ % RC::DS] = st.gb[clamp64(t - RC::DS, 0, st.T - 1)];
      if (live && st.dxb) st.dxb[t] = convert<S>(lam);
    }
    if constexpr (DA == DA_PARTIAL) {
#pragma unroll
      for (int i = 1; i <= NP; ++i)
        st.dacc[i - 1] = Num<W>::fms(st.dacc[i - 1], lam, convert<W>(st.yr[(j - i) & (NP - 1)]));
      // slide the window: slot (t-1) mod NP now receives y_{t-1-NP}; its stream
      // slot is consumed here, before it is reloaded
      if constexpr (VEC) {
        // y_{t-1-NP} sits at element (2 - JJ - NP) mod 4 of block slot
        // (QB - 1 + floor((2 - JJ - NP) / 4)) mod QB: tb = 4m + 3 with m = QB - 1
        // mod QB at every unrolled block (chunk ends are multiples of 16)
        constexpr int QB = RC::DS / 4;
        constexpr int ye = posmod(2 - JJ - NP, 4);
        constexpr int yslot = posmod(QB - 1 + floordiv4(2 - JJ - NP), QB);
        st.yr[(j - 1) & (NP - 1)] = v4get<ye>(st.yq[yslot]);
        if constexpr (ye == 0) st.yq[yslot] = ywin_ti(st, t - 1 - NP - RC::DS);
      } else {
        const bool ynew_ok = yw_ok(st.zib, a.n, st.T, t - 1 - NP);
        st.yr[(j - 1) & (NP - 1)] = ynew_ok ? Num<S>::conj(st.ys[JJ % RC::DS]) : Num<S>::zero();
        st.ys[JJ % RC::DS] = *yw_addr(st.yb, st.zib, a.n, st.T, t - RC::DS - 1 - NP);
      }
    }
    pin_step();
    bwd_steps_ti<S, NP, DA, VEC, JJ + 1>(a, st, tb);
  }
}

// Adjoint rerun, time-invariant: one lane per chunk walks its samples backwards
// from the carry, writes dx and (DA_PARTIAL) the chunk's partial of da; chunk 0
// writes dzi.
template <typename S, int NP, int DA, bool VEC>
__global__ __launch_bounds__(256) void bwd_rerun_ti(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cid >= a.B * a.P) return;
  const int64_t b = cid / a.P;
  const int64_t c = cid - b * a.P;
  const int64_t t0 = c * a.L;
  BwdLaneTI<S, NP, VEC> st;
  st.gb = a.x + b * a.T;  // upstream gradient
  st.yb = a.y + b * a.T;
  st.zib = a.zi ? a.zi + b * a.n : nullptr;
  st.dxb = a.dx ? a.dx + b * a.T : nullptr;
  st.T = a.T;
  const int64_t tstart = t0 + a.L - 1;  // first processed (possibly virtual) time
#pragma unroll
  for (int k = 0; k < RC::DS; ++k) {
    const int64_t t = tstart - k;
    if constexpr (!VEC) st.gs[k] = opaque_v(st.gb[clamp64(t, 0, a.T - 1)]);
    if constexpr (DA == DA_PARTIAL && !VEC) st.ys[k] = opaque_v(*yw_addr(st.yb, st.zib, a.n, a.T, t - 1 - NP));
  }
  if constexpr (DA == DA_PARTIAL && VEC) {
    // the y blocks that hold positions tstart-1-NP and the DS - 4 below it (the
    // slot of block u0 - 4k is a constant once the loop is unrolled)
    constexpr int QB = RC::DS / 4;
    st.ypadb = reinterpret_cast<const float*>(a.ypad) + b * ti_ypad(NP);
    const int64_t u0 = tstart - 1 - NP - posmod(2 - NP, 4);
#pragma unroll
    for (int k = 0; k < QB; ++k)
      st.yq[posmod(QB - 1 + floordiv4(2 - NP) - k, QB)] = opaque_v(ywin_ti(st, u0 - 4 * k));
  }
  if constexpr (VEC) {
#pragma unroll
    for (int q = 0; q < RC::DS / 4; ++q) st.gq[q] = opaque_v(ld4(st.gb, clamp64(tstart - 3 - 4 * q, 0, a.T - 4)));
  }
  load_coefs_w<S, NP>(st.cw, a, a.A + b * a.a_bstride);
#pragma unroll
  for (int i = 0; i < NP; ++i) st.cw[i] = Num<W>::conj(st.cw[i]);
  if constexpr (DA == DA_PARTIAL) {
#pragma unroll
    for (int i = 1; i <= NP; ++i) {
      const int64_t u = tstart - i;
      const S v = *yw_addr(st.yb, st.zib, a.n, a.T, u);
      st.yr[(NP - 1 - i + NP) & (NP - 1)] = yw_ok(st.zib, a.n, a.T, u) ? Num<S>::conj(v) : Num<S>::zero();
    }
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) st.dacc[i] = Num<W>::zero();
  const W* cr = a.carry + cid * NP;
#pragma unroll
  for (int r = 0; r < NP; ++r) st.ring[NP - 1 - r] = opaque_v(cr[r]);
  for (int jb = 0; jb < a.L; jb += RC::K) bwd_steps_ti<S, NP, DA, VEC, 0>(a, st, tstart - jb);
  if constexpr (DA == DA_PARTIAL) {
    W* dp = a.dapart + cid * NP;
#pragma unroll
    for (int i = 0; i < NP; ++i) dp[i] = a.neg ? -st.dacc[i] : st.dacc[i];
  }
  if (c == 0 && a.dzi) {
    // pending sums for positions -1-r sit in slots NP-1-r
#pragma unroll
    for (int r = 0; r < NP; ++r)
      if (r < a.n) a.dzi[b * a.n + r] = convert<S>(st.ring[NP - 1 - r]);
  }
}

// Phase 3 (reruns), general: one lane per chunk, loading each step's row and
// sample as the step comes. Time-varying rows do not fit a prefetch ring in the
// lane's registers at the higher orders. Ragged batches, other dtypes, orders
// and alignments, and a full dA of time-invariant filters take these kernels.
//
// Up to order 4 (BL) the steps are branch-free: loads are unconditional, at
// times clamped into the chunk, so the compiler may issue a step's loads ahead
// of the previous step's arithmetic (a load under a branch is waited for at
// the branch's end); steps past the end compute on clamped data, and their
// stores go to the lane's own carry slot (already read, and rewritten by the
// next carry pass) instead of being skipped (a store under a divergent branch
// also makes the wait-count pass drain the loads in flight). At orders 8 to 32
// that form measured slower in the ragged batch (more registers, the row's
// loads interleaved with its products), so those steps keep the branch.
//
// Forward: one lane per chunk re-filters from its carry and writes y.
// The recurrence runs in the wide type (float64 for float32 data), so the only
// single-precision rounding of an output is its final store.
template <typename S, int NP>
__global__ __launch_bounds__(256) void fwd_rerun(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cid >= (a.rg ? a.G : a.B * a.P)) return;
  const int64_t b = chunk_signal(a, cid);
  const Span sp = sig_span(a, b);
  const int64_t c = cid - sp.cbase;
  const int64_t t0 = c * a.L;
  const int64_t rem = sp.T - t0;
  const int len = rem < a.L ? (int)rem : a.L;
  W ring[NP];
  const W* cr = a.carry + cid * NP;
#pragma unroll
  for (int r = 0; r < NP; ++r) ring[NP - 1 - r] = cr[r];
  const S* xb = a.x + sp.xo;
  const S* Ab = a.A + sp.ao;
  S* yb = a.y + sp.xo;
  constexpr bool BL = NP <= 4;
  // steps past the end (the last chunk's tail) sit at the clamped time T-1 and
  // store the output of T-1 there again, so that no store waits on a branch and
  // the chunk carries stay as the carry pass left them (an adjoint that rebuilds
  // y reads them)
  S ylast = Num<S>::zero();
  for (int jb = 0; jb < len; jb += NP) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      if constexpr (BL) {
        const bool live = jb + j < len;
        const int64_t t = t0 + (live ? jb + j : len - 1);
        S row[NP];
        load_row_u<S, NP>(row, a, Ab + t * sp.at);
        W acc = convert<W>(xb[t]);
#pragma unroll
        for (int i = NP; i >= 1; --i)
          acc = Num<W>::fms(acc, convert<W>(row[i - 1]), ring[(j - i) & (NP - 1)]);
        ring[j] = acc;
        const S yv = convert<S>(acc);
        ylast = live ? yv : ylast;
        yb[t] = ylast;
      } else if (jb + j < len) {
        const int64_t t = t0 + jb + j;
        S row[NP];
        load_row<S, NP>(row, a, Ab + t * sp.at);
        W acc = convert<W>(xb[t]);
#pragma unroll
        for (int i = NP; i >= 1; --i)
          acc = Num<W>::fms(acc, convert<W>(row[i - 1]), ring[(j - i) & (NP - 1)]);
        ring[j] = acc;
        yb[t] = convert<S>(acc);
      }
    }
  }
}

// y_u of one signal for the adjoint's window: zi for u < 0 (k < n), zero past T.
template <typename S>
__device__ __forceinline__ S yval_at(const S* yb, const S* zib, int n, int64_t T, int64_t u) {
  if (u >= 0) return (u < T) ? yb[u] : Num<S>::zero();
  const int64_t k = -u - 1;
  return (zib && k < n) ? zib[k] : Num<S>::zero();
}

// One lane's adjoint state and streams (see bwd_rerun).
template <typename S, int NP>
struct BwdLane {
  using W = typename Wide<S>::type;
  W ring[NP];  // pending sums, slot u mod NP for position u
  S yr[NP];    // conj(y) window: slot u mod NP holds y_u for u in [t-NP, t-1] (data type, exact)
  W dacc[NP];  // time-invariant da partials
  const S* gb;
  const S* Ab;
  const S* yb;
  const S* zib;
  S* dxb;
  S* dAb;
  S* sink;  // the lane's carry slot: destination of the stores of steps past T
  int64_t at, T;
  bool need_da;
};

// One adjoint step under a branch (orders 8 to 32; see the note above fwd_rerun).
template <typename S, int NP, int JJ>
__device__ __forceinline__ void bwd_step_branchy(const ChunkArgs<S>& a, BwdLane<S, NP>& st, int64_t t) {
  using W = typename Wide<S>::type;
  constexpr int j = NP - 1 - JJ;
  if (t < st.T) {
    S row[NP];
    load_row<S, NP>(row, a, st.Ab + t * st.at);
    const W lam = convert<W>(st.gb[t]) + st.ring[j];
    st.ring[j] = Num<W>::zero();
#pragma unroll
    for (int i = 1; i <= NP; ++i)
      st.ring[(j - i) & (NP - 1)] =
          Num<W>::fms(st.ring[(j - i) & (NP - 1)], Num<W>::conj(convert<W>(row[i - 1])), lam);
    if (st.dxb) st.dxb[t] = convert<S>(lam);
    if (st.need_da) {
      if (st.dAb) {
        S* drow = st.dAb + t * a.n;
        S d[NP];
#pragma unroll
        for (int i = 1; i <= NP; ++i) d[i - 1] = coef(a, convert<S>(-(lam * convert<W>(st.yr[(j - i) & (NP - 1)]))));
        bool vec = false;
        if constexpr (std::is_same<S, float>::value && NP % 4 == 0) {
          if (a.n == NP && (reinterpret_cast<uintptr_t>(drow) & 15) == 0) {
            float4* q = reinterpret_cast<float4*>(drow);
#pragma unroll
            for (int k = 0; k < NP / 4; ++k) q[k] = make_float4(d[4 * k], d[4 * k + 1], d[4 * k + 2], d[4 * k + 3]);
            vec = true;
          }
        }
        if (!vec) {
#pragma unroll
          for (int i = 1; i <= NP; ++i)
            if (i <= a.n) drow[i - 1] = d[i - 1];
        }
      } else {
#pragma unroll
        for (int i = 1; i <= NP; ++i)
          st.dacc[i - 1] = Num<W>::fms(st.dacc[i - 1], lam, convert<W>(st.yr[(j - i) & (NP - 1)]));
      }
    }
  }
  if (st.need_da) {
    // slide the window: slot (t-1) mod NP now receives y_{t-1-NP}
    st.yr[(j - 1) & (NP - 1)] = Num<S>::conj(yval_at(st.yb, st.zib, a.n, st.T, t - 1 - NP));
  }
}

// Steps t = tb - JJ for JJ = 0..NP-1 (t mod NP = NP-1-JJ), unrolled by template
// recursion so that every ring index is a compile-time constant (a runtime index
// would put the float64 ring in scratch memory).
template <typename S, int NP, int JJ>
__device__ __forceinline__ void bwd_steps(const ChunkArgs<S>& a, BwdLane<S, NP>& st, int64_t tb) {
  using W = typename Wide<S>::type;
  if constexpr (JJ < NP) {
    constexpr int j = NP - 1 - JJ;
    const int64_t t = tb - JJ;
    if constexpr (NP > 4) {
      bwd_step_branchy<S, NP, JJ>(a, st, t);
    } else {
      // steps past the end of the signal (the last chunk's tail) have lambda = 0
      // and leave the pending sums as they are
      const bool live = t < st.T;
      const int64_t tc = live ? t : st.T - 1;
      S row[NP];
      load_row_u<S, NP>(row, a, st.Ab + tc * st.at);
      const W lam0 = convert<W>(st.gb[tc]) + st.ring[j];
      const W lam = live ? lam0 : Num<W>::zero();
      st.ring[j] = live ? Num<W>::zero() : st.ring[j];
#pragma unroll
      for (int i = 1; i <= NP; ++i)
        st.ring[(j - i) & (NP - 1)] =
            Num<W>::fms(st.ring[(j - i) & (NP - 1)], Num<W>::conj(convert<W>(row[i - 1])), lam);
      if (st.dxb) *(live ? st.dxb + t : st.sink) = convert<S>(lam);
      if (st.dAb) {
        S* drow = live ? st.dAb + t * a.n : st.sink;
#pragma unroll
        for (int i = 1; i <= NP; ++i)
          if (i <= a.n) drow[i - 1] = coef(a, convert<S>(-(lam * convert<W>(st.yr[(j - i) & (NP - 1)]))));
      } else if (st.need_da) {
#pragma unroll
        for (int i = 1; i <= NP; ++i)
          st.dacc[i - 1] = Num<W>::fms(st.dacc[i - 1], lam, convert<W>(st.yr[(j - i) & (NP - 1)]));
      }
      // slide the window: slot (t-1) mod NP now receives y_{t-1-NP} (read at an
      // always-valid address, selected afterwards; unused without da)
      const int64_t u = t - 1 - NP;
      const S yv = *yw_addr(st.yb, st.zib, a.n, st.T, u);
      st.yr[(j - 1) & (NP - 1)] = yw_ok(st.zib, a.n, st.T, u) ? Num<S>::conj(yv) : Num<S>::zero();
    }
    bwd_steps<S, NP, JJ + 1>(a, st, tb);
  }
}

// Phase 3 (adjoint): one lane per chunk walks its samples backwards from the carry,
// writes dx and dA (time-varying) or a per-chunk partial of da (time-invariant),
// and chunk 0 writes dzi. Arithmetic in the wide type, as in the forward rerun.
template <typename S, int NP>
__global__ __launch_bounds__(256) void bwd_rerun(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cid >= (a.rg ? a.G : a.B * a.P)) return;
  const int64_t b = chunk_signal(a, cid);
  const Span sp = sig_span(a, b);
  const int64_t c = cid - sp.cbase;
  const int64_t t0 = c * a.L;
  BwdLane<S, NP> st;
  const W* cr = a.carry + cid * NP;
#pragma unroll
  for (int r = 0; r < NP; ++r) st.ring[NP - 1 - r] = cr[r];
  st.gb = a.x + sp.xo;  // upstream gradient
  st.Ab = a.A + sp.ao;
  st.yb = a.y + sp.xo;
  st.zib = a.zi ? a.zi + sp.zo : nullptr;
  st.dxb = a.dx ? a.dx + sp.xo : nullptr;
  st.dAb = a.dA ? a.dA + sp.dao : nullptr;
  st.sink = reinterpret_cast<S*>(a.carry + cid * NP);
  st.at = sp.at;
  st.T = sp.T;
  st.need_da = a.dA != nullptr || a.dapart != nullptr;
  const int64_t tstart = t0 + a.L - 1;  // first processed (possibly virtual) time
  if (st.need_da) {
#pragma unroll
    for (int i = 1; i <= NP; ++i)
      st.yr[(NP - 1 - i + NP) & (NP - 1)] = Num<S>::conj(yval_at(st.yb, st.zib, a.n, st.T, tstart - i));
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) st.dacc[i] = Num<W>::zero();
  for (int jb = 0; jb < a.L; jb += NP) bwd_steps<S, NP, 0>(a, st, tstart - jb);
  if (a.dapart) {
    W* dp = a.dapart + cid * NP;
#pragma unroll
    for (int i = 0; i < NP; ++i) dp[i] = a.neg ? -st.dacc[i] : st.dacc[i];
  }
  if (c == 0 && a.dzi) {
    // pending sums for positions -1-r sit in slots NP-1-r
#pragma unroll
    for (int r = 0; r < NP; ++r)
      if (r < a.n) a.dzi[sp.zo + r] = convert<S>(st.ring[NP - 1 - r]);
  }
}

// ----------------------------------------------------------------------------
// Adjoint rerun without a saved y (time-varying, uniform batch, orders up to 4,
// real data): each lane rebuilds its chunk's y from the forward's chunk carry
// (a.fcarry) and the forward input (a.rx). Pass 1 walks the chunk forwards and
// keeps the wide-type state at the start of every 16-sample sub-block in LDS;
// pass 2 takes the sub-blocks last to first, re-filters each one's 16 outputs
// into registers (the same operations as fwd_rerun, so the same values) and
// runs the adjoint steps of bwd_steps over them. 64-lane workgroups; chunks of
// at most RC_LMAX samples.
constexpr int RC_SB = 16;
constexpr int RC_LMAX = 256;

template <typename S, int NP, int JJ>
__device__ __forceinline__ void rc_fwd_steps(typename Wide<S>::type (&f)[NP], S (&yl)[RC_SB + NP],
                                             S (&rows)[RC_SB][NP], const ChunkArgs<S>& a, const S* xb,
                                             const S* Ab, int64_t tq, int64_t T) {
  using W = typename Wide<S>::type;
  if constexpr (JJ < RC_SB) {
    constexpr int j = JJ % NP;  // tq is a multiple of NP
    const int64_t t = tq + JJ < T ? tq + JJ : T - 1;
    load_row_u<S, NP>(rows[JJ], a, Ab + t * a.a_tstride);
    W acc = convert<W>(xb[t]);
#pragma unroll
    for (int i = NP; i >= 1; --i) acc = Num<W>::fms(acc, convert<W>(rows[JJ][i - 1]), f[(j - i) & (NP - 1)]);
    f[j] = acc;
    yl[NP + JJ] = convert<S>(acc);
    rc_fwd_steps<S, NP, JJ + 1>(f, yl, rows, a, xb, Ab, tq, T);
  }
}

template <typename S, int NP, int JJ>
__device__ __forceinline__ void rc_adj_steps(typename Wide<S>::type (&ring)[NP], const S (&yl)[RC_SB + NP],
                                             const S (&rows)[RC_SB][NP], const ChunkArgs<S>& a, const S* gb,
                                             S* dxb, S* dAb, S* sink, int64_t tq, int64_t T) {
  using W = typename Wide<S>::type;
  if constexpr (JJ < RC_SB) {
    constexpr int s = RC_SB - 1 - JJ;  // sub-block position, last first
    constexpr int j = s % NP;
    const int64_t t = tq + s;
    const bool live = t < T;
    const int64_t tc = live ? t : T - 1;
    const S (&row)[NP] = rows[s];  // the rows the rebuild of y just loaded
    const W lam0 = convert<W>(gb[tc]) + ring[j];
    const W lam = live ? lam0 : Num<W>::zero();
    ring[j] = live ? Num<W>::zero() : ring[j];
#pragma unroll
    for (int i = 1; i <= NP; ++i)
      ring[(j - i) & (NP - 1)] = Num<W>::fms(ring[(j - i) & (NP - 1)], Num<W>::conj(convert<W>(row[i - 1])), lam);
    if (dxb) *(live ? dxb + t : sink) = convert<S>(lam);
    if (dAb) {
      S* drow = live ? dAb + t * a.n : sink;
#pragma unroll
      for (int i = 1; i <= NP; ++i)
        if (i <= a.n) drow[i - 1] = coef(a, convert<S>(-(lam * convert<W>(yl[NP + s - i]))));
    }
    rc_adj_steps<S, NP, JJ + 1>(ring, yl, rows, a, gb, dxb, dAb, sink, tq, T);
  }
}

template <typename S, int NP>
__global__ __launch_bounds__(64) void bwd_rerun_rc(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  __shared__ W sbs[(RC_LMAX / RC_SB) * NP * 64];
  const int lane = threadIdx.x;
  const int64_t cid = (int64_t)blockIdx.x * 64 + lane;
  if (cid >= a.B * a.P) return;
  const int64_t b = cid / a.P;
  const int64_t c = cid - b * a.P;
  const int64_t t0 = c * a.L;
  const int64_t T = a.T;
  const int nq = a.L / RC_SB;
  const S* xb = a.rx + b * T;
  const S* gb = a.x + b * T;  // upstream gradient
  const S* Ab = a.A + b * a.a_bstride;
  S* dxb = a.dx ? a.dx + b * T : nullptr;
  S* dAb = a.dA ? a.dA + b * a.a_bstride : nullptr;
  S* sink = reinterpret_cast<S*>(a.carry + cid * NP);  // already read; rewritten by the next carry pass
  // pass 1: the forward state at each sub-block start (slot u mod NP holds y_u)
  W f[NP];
  const W* fc = a.fcarry + cid * NP;
#pragma unroll
  for (int r = 0; r < NP; ++r) f[NP - 1 - r] = fc[r];
  S yl[RC_SB + NP];
  S rows[RC_SB][NP];
  for (int q = 0; q < nq; ++q) {
#pragma unroll
    for (int r = 0; r < NP; ++r) sbs[(q * NP + r) * 64 + lane] = f[r];
    rc_fwd_steps<S, NP, 0>(f, yl, rows, a, xb, Ab, t0 + (int64_t)q * RC_SB, T);
  }
  // pass 2: adjoint over the sub-blocks, last first, from the adjoint carry
  W ring[NP];
  const W* cr = a.carry + cid * NP;
#pragma unroll
  for (int r = 0; r < NP; ++r) ring[NP - 1 - r] = cr[r];
  for (int q = nq - 1; q >= 0; --q) {
    const int64_t tq = t0 + (int64_t)q * RC_SB;
#pragma unroll
    for (int r = 0; r < NP; ++r) f[r] = sbs[(q * NP + r) * 64 + lane];
    // y_{tq-NP+k} = f[k] before the sub-block (y_{-1-k} = zi[k] for chunk 0: the
    // forward's first carry), then the sub-block's own outputs
#pragma unroll
    for (int k = 0; k < NP; ++k) yl[k] = convert<S>(f[k]);
    rc_fwd_steps<S, NP, 0>(f, yl, rows, a, xb, Ab, tq, T);
    rc_adj_steps<S, NP, 0>(ring, yl, rows, a, gb, dxb, dAb, sink, tq, T);
  }
  if (c == 0 && a.dzi) {
#pragma unroll
    for (int r = 0; r < NP; ++r)
      if (r < a.n) a.dzi[b * a.n + r] = convert<S>(ring[NP - 1 - r]);
  }
}

// ----------------------------------------------------------------------------
// Reruns with float4 streams, time-varying float32 filters of order 2, 4 or 8
// (n == NP) in a uniform batch with T a multiple of 4 and 16-byte aligned
// streams. One lane per chunk, as in fwd_rerun / bwd_rerun, but every stream
// moves in 4-step blocks: the block's NP coefficient rows are NP float4, and x,
// g, y, dx are one float4 each. The blocks are prefetched DS steps ahead in
// rings whose slots are consumed before they are reloaded, all loads
// unconditional at clamped block starts. A lane's 4-byte accesses are L2
// requests of their own; the per-step kernels issue NP + 2 (forward) or 2NP + 4
// (adjoint) of them per sample, these kernels about a quarter.
template <int NP>
struct Tv4 {
  static constexpr int DS = NP <= 4 ? 16 : 8;  // prefetch distance in steps
  static constexpr int QB = DS / 4;            // ring slots (4-step blocks)
  static constexpr int K = 16;                 // steps per unrolled block
};

// The 4*M floats of M float4 as one array (indices fold to registers once the
// callers' loops are unrolled).
template <int M>
__device__ __forceinline__ void flat4(const float4 (&v)[M], float (&f)[4 * M]) {
#pragma unroll
  for (int k = 0; k < M; ++k) {
    f[4 * k] = v[k].x;
    f[4 * k + 1] = v[k].y;
    f[4 * k + 2] = v[k].z;
    f[4 * k + 3] = v[k].w;
  }
}

template <int NP>
struct FwdLaneTv4 {
  using C = Tv4<NP>;
  double ring[NP];            // y_u in slot u mod NP
  float4 aq[C::QB][NP];       // rows of a 4-step block: 4*NP floats = NP float4
  float4 xq[C::QB];
  float4 yq;
  const float* xb;
  const float* Ab;
  float* yb;
  int64_t T, tend;
};

template <int NP, int JJ>
__device__ __forceinline__ void fwd_steps_tv4(FwdLaneTv4<NP>& st, int64_t tb) {
  using C = Tv4<NP>;
  if constexpr (JJ < C::K) {
    constexpr int j = JJ % NP;
    constexpr int q = (JJ % C::DS) / 4;
    constexpr int r = JJ & 3;
    const int64_t t = tb + JJ;
    float cf[4 * NP];
    flat4<NP>(st.aq[q], cf);
    double acc = (double)v4get<r>(st.xq[q]);
#pragma unroll
    for (int i = NP; i >= 1; --i) acc -= (double)cf[r * NP + i - 1] * st.ring[(j - i) & (NP - 1)];
    if constexpr (r == 3) {
      // the block's last step: its slots take the block DS steps ahead
      const int64_t u = min(t - 3 + C::DS, st.T - 4);
      const float4* ap = reinterpret_cast<const float4*>(st.Ab + u * NP);
#pragma unroll
      for (int k = 0; k < NP; ++k) st.aq[q][k] = ap[k];
      st.xq[q] = ld4(st.xb, u);
    }
    pin_step();
    st.ring[j] = acc;
    v4set<r>(st.yq, (float)acc);
    if constexpr (r == 3)
      if (t < st.tend) *reinterpret_cast<float4*>(st.yb + t - 3) = st.yq;
    fwd_steps_tv4<NP, JJ + 1>(st, tb);
  }
}

template <int NP>
__global__ __launch_bounds__(256) void fwd_rerun_tv4(ChunkArgs<float> a) {
  using C = Tv4<NP>;
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cid >= a.B * a.P) return;
  const int64_t b = cid / a.P;
  const int64_t c = cid - b * a.P;
  const int64_t t0 = c * a.L;
  const int64_t rem = a.T - t0;
  const int len = rem < a.L ? (int)rem : a.L;
  FwdLaneTv4<NP> st;
  st.xb = a.x + b * a.T;
  st.Ab = a.A + b * a.a_bstride;
  st.yb = a.y + b * a.T;
  st.T = a.T;
  st.tend = t0 + len;
#pragma unroll
  for (int q = 0; q < C::QB; ++q) {
    const int64_t u = min(t0 + 4 * q, a.T - 4);
    const float4* ap = reinterpret_cast<const float4*>(st.Ab + u * NP);
#pragma unroll
    for (int k = 0; k < NP; ++k) st.aq[q][k] = opaque_v(ap[k]);
    st.xq[q] = opaque_v(ld4(st.xb, u));
  }
  const double* cr = a.carry + cid * NP;
#pragma unroll
  for (int r = 0; r < NP; ++r) st.ring[NP - 1 - r] = opaque_v(cr[r]);
  for (int jb = 0; jb < len; jb += C::K) fwd_steps_tv4<NP, 0>(st, t0 + jb);
}

template <int NP>
struct BwdLaneTv4 {
  using C = Tv4<NP>;
  double p[NP];          // pending sums, slot u mod NP for position u
  float yr[NP];          // y window: slot u mod NP holds y_u for u in [t-NP, t-1]
  float zr[NP];          // y_{-1-k} = zi[k] (zero past the order or without zi)
  float4 aq[C::QB][NP];  // coefficient rows of 4-step blocks
  float4 gq[C::QB];
  float4 yq[C::QB];      // y blocks of the window input (positions t-1-NP)
  float4 dq;             // lambda of the current block
  float daf[4 * NP];     // dA rows of the current block
  const float* gb;
  const float* Ab;
  const float* yb;
  float* dxb;
  float* dAb;
  int64_t T;
};

// y[u .. u+3] as the window reads it: y_{-1-k} = zi[k] below 0 (chunk 0 only),
// clamped into the signal above (values past T meet only steps past T).
template <int NP>
__device__ __forceinline__ float4 yblock(const BwdLaneTv4<NP>& st, int64_t u) {
  float4 v = ld4(st.yb, clamp64(u, 0, st.T - 4));
  float c[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int64_t k = -(u + m) - 1;  // zi index when u + m < 0
    float z = 0.f;
#pragma unroll
    for (int kk = 0; kk < NP; ++kk) z = (k == kk) ? st.zr[kk] : z;
    c[m] = (u + m < 0) ? z : c[m];
  }
  return make_float4(c[0], c[1], c[2], c[3]);
}

// Steps t = tb - JJ (tb is 3 mod 4 and NP-1 mod NP). Block slot q holds
// positions [u, u+3] with u = t - e, e = 3 - JJ % 4. The window input
// y_{t-1-NP} sits at element (2 - JJ - NP) mod 4 of y slot
// (QB - 1 + floor((2 - JJ - NP) / 4)) mod QB (the slot index of the block that
// holds it: tb = 4m + 3 with m = QB - 1 mod QB at every unrolled block).
template <int NP, int JJ>
__device__ __forceinline__ void bwd_steps_tv4(BwdLaneTv4<NP>& st, int64_t tb) {
  using C = Tv4<NP>;
  if constexpr (JJ < C::K) {
    constexpr int j = NP - 1 - JJ % NP;
    constexpr int q = (JJ % C::DS) / 4;
    constexpr int e = 3 - (JJ & 3);
    constexpr int ye = posmod(2 - JJ - NP, 4);
    constexpr int yslot = posmod(C::QB - 1 + floordiv4(2 - JJ - NP), C::QB);
    const int64_t t = tb - JJ;
    const bool live = t < st.T;
    const double lam0 = (double)v4get<e>(st.gq[q]) + st.p[j];
    const double lam = live ? lam0 : 0.0;
    st.p[j] = live ? 0.0 : st.p[j];
    float cf[4 * NP];
    flat4<NP>(st.aq[q], cf);
#pragma unroll
    for (int i = 1; i <= NP; ++i) st.p[(j - i) & (NP - 1)] -= (double)cf[e * NP + i - 1] * lam;
    v4set<e>(st.dq, (float)lam);
#pragma unroll
    for (int i = 1; i <= NP; ++i) st.daf[e * NP + i - 1] = (float)(-(lam * (double)st.yr[(j - i) & (NP - 1)]));
    // slide the window: slot (t-1) mod NP now receives y_{t-1-NP}
    st.yr[(j - 1) & (NP - 1)] = v4get<ye>(st.yq[yslot]);
    if constexpr (ye == 0) st.yq[yslot] = yblock<NP>(st, t - 1 - NP - C::DS);
    if constexpr (e == 0) {
      // u = t: the block is complete; its slots take the block DS steps below
      const int64_t u = clamp64(t - C::DS, 0, st.T - 4);
      const float4* ap = reinterpret_cast<const float4*>(st.Ab + u * NP);
#pragma unroll
      for (int k = 0; k < NP; ++k) st.aq[q][k] = ap[k];
      st.gq[q] = ld4(st.gb, u);
      // T is a multiple of 4: the block is all live or all past T
      if (live) {
        if (st.dxb) *reinterpret_cast<float4*>(st.dxb + t) = st.dq;
        if (st.dAb) {
          float4* dp = reinterpret_cast<float4*>(st.dAb + t * NP);
#pragma unroll
          for (int k = 0; k < NP; ++k)
            dp[k] = make_float4(st.daf[4 * k], st.daf[4 * k + 1], st.daf[4 * k + 2], st.daf[4 * k + 3]);
        }
      }
    }
    pin_step();
    bwd_steps_tv4<NP, JJ + 1>(st, tb);
  }
}

template <int NP>
__global__ __launch_bounds__(256) void bwd_rerun_tv4(ChunkArgs<float> a) {
  using C = Tv4<NP>;
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cid >= a.B * a.P) return;
  const int64_t b = cid / a.P;
  const int64_t c = cid - b * a.P;
  const int64_t t0 = c * a.L;
  BwdLaneTv4<NP> st;
  st.gb = a.x + b * a.T;  // upstream gradient
  st.Ab = a.A + b * a.a_bstride;
  st.yb = a.y + b * a.T;
  st.dxb = a.dx ? a.dx + b * a.T : nullptr;
  st.dAb = a.dA ? a.dA + b * a.T * NP : nullptr;
  st.T = a.T;
  const float* zib = a.zi ? a.zi + b * a.n : nullptr;
#pragma unroll
  for (int k = 0; k < NP; ++k) st.zr[k] = zib ? zib[k] : 0.f;
  const int64_t tstart = t0 + a.L - 1;  // 3 mod 4 and NP-1 mod NP
#pragma unroll
  for (int q = 0; q < C::QB; ++q) {
    const int64_t u = clamp64(tstart - 3 - 4 * q, 0, a.T - 4);
    const float4* ap = reinterpret_cast<const float4*>(st.Ab + u * NP);
#pragma unroll
    for (int k = 0; k < NP; ++k) st.aq[q][k] = opaque_v(ap[k]);
    st.gq[q] = opaque_v(ld4(st.gb, u));
  }
  // the y blocks that hold positions tstart-1-NP and the DS - 4 below it
  {
    const int64_t pf = tstart - 1 - NP;
    const int64_t u0 = pf - posmod(2 - NP, 4);  // block start of pf
    // (the slot of block u0 - 4k is a constant once the loop is unrolled)
#pragma unroll
    for (int k = 0; k < C::QB; ++k)
      st.yq[posmod(C::QB - 1 + floordiv4(2 - NP) - k, C::QB)] = opaque_v(yblock<NP>(st, u0 - 4 * k));
  }
  // window y_{tstart-1} .. y_{tstart-NP}
#pragma unroll
  for (int i = 1; i <= NP; ++i) {
    const int64_t u = tstart - i;
    const float v = *yw_addr(st.yb, zib, a.n, a.T, u);
    st.yr[(NP - 1 - i + NP) & (NP - 1)] = yw_ok(zib, a.n, a.T, u) ? v : 0.f;
  }
  const double* cr = a.carry + cid * NP;
#pragma unroll
  for (int r = 0; r < NP; ++r) st.p[NP - 1 - r] = opaque_v(cr[r]);
  for (int jb = 0; jb < a.L; jb += C::K) bwd_steps_tv4<NP, 0>(st, tstart - jb);
  if (c == 0 && a.dzi) {
    // pending sums for positions -1-r sit in slots NP-1-r
#pragma unroll
    for (int r = 0; r < NP; ++r) a.dzi[b * a.n + r] = (float)st.p[NP - 1 - r];
  }
}

// Time-invariant da: deterministic sum of the per-chunk partials (wide type).
template <typename S, int NP>
__global__ __launch_bounds__(256) void ti_da_reduce(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.B * a.n) return;
  const int64_t b = idx / a.n;
  const int i = (int)(idx - b * a.n);
  W acc = Num<W>::zero();
  const W* dp = a.dapart + b * a.P * NP + i;
  // unrolled so that 16 loads are in flight; the sum keeps chunk order
#pragma unroll 16
  for (int64_t c = 0; c < a.P; ++c) acc = acc + dp[c * NP];
  a.da_ti[b * a.n + i] = convert<S>(acc);
}

// ----------------------------------------------------------------------------
// Time-invariant coefficients (one row a[n] per signal, a_tstride == 0). Every
// full chunk of a signal has the same transition matrix Phi = C^L (C the
// companion matrix of a), so the summary phase only needs each chunk's
// zero-state response (O(n) per sample, one lane per chunk), and the carry
// kernel builds Phi once per signal from the n basis responses over L steps.
// The adjoint's chunk map is Phi^H: the conjugated, transposed recurrence is the
// adjoint of the forward one. Summaries hold only the NP-vector per chunk.

template <typename S, int NP>
__device__ __forceinline__ void load_coefs(typename Wide<S>::type (&co)[NP], const ChunkArgs<S>& a,
                                           int64_t b, bool conj) {
  using W = typename Wide<S>::type;
  const S* arow = a.A + b * a.a_bstride;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const W v = i < a.n ? convert<W>(coef(a, arow[i])) : Num<W>::zero();
    co[i] = conj ? Num<W>::conj(v) : v;
  }
}

// Zero-state response of one chunk, both directions in the pending-sum form: the
// step's NP products are independent, and only the first feeds the next step.
// Forward step t: y = x_t + p(t), then p(t+i) -= a_i y (i = 1..NP) and y joins
// the history window yh. Adjoint step t (descending): lambda = g_t + p(t), then
// p(t-i) -= conj(a_i) lambda. The input stream is prefetched DS steps ahead in a
// ring whose slots are consumed before they are reloaded; with VEC (float32,
// T a multiple of 4, input 16-byte aligned) as float4 blocks.
template <typename S, int NP, bool VEC>
struct ZsumLane {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  W p[NP];                          // pending sums, slot u mod NP for position u
  W yh[NP];                         // forward: y_u in slot u mod NP
  W co[NP];                         // coefficients (conjugated for the adjoint)
  S xs[VEC ? 1 : RC::DS];           // input sample in slot JJ mod DS
  float4 xq[VEC ? RC::DS / 4 : 1];  // VEC: input blocks
  const S* inp;
  int64_t T;
};

template <typename S, int NP, bool VEC, bool FWD, int JJ>
__device__ __forceinline__ void zsum_steps(ZsumLane<S, NP, VEC>& st, int64_t tb) {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  if constexpr (JJ < RC::K) {
    // forward: t = tb + JJ (tb a multiple of NP and of 4); adjoint: t = tb - JJ
    // (tb is NP-1 mod NP and 3 mod 4)
    constexpr int j = FWD ? JJ % NP : NP - 1 - JJ % NP;
    constexpr int comp = FWD ? (JJ & 3) : 3 - (JJ & 3);
    const int64_t t = FWD ? tb + JJ : tb - JJ;
    W v;
    if constexpr (VEC)
      v = convert<W>(v4get<comp>(st.xq[(JJ % RC::DS) / 4]));
    else
      v = convert<W>(st.xs[JJ % RC::DS]);
    if constexpr (VEC) {
      if constexpr ((JJ & 3) == 3)
        st.xq[(JJ % RC::DS) / 4] =
            ld4(st.inp, FWD ? min(t - 3 + RC::DS, st.T - 4) : clamp64(t - RC::DS, 0, st.T - 4));
    } else {
      st.xs[JJ % RC::DS] = st.inp[FWD ? min(t + RC::DS, st.T - 1) : clamp64(t - RC::DS, 0, st.T - 1)];
    }
    if constexpr (FWD) {
      // chunks 0..P-2 are whole: t < T throughout
      const W y = v + st.p[j];
      st.p[j] = Num<W>::zero();
#pragma unroll
      for (int i = 1; i <= NP; ++i) st.p[(j + i) & (NP - 1)] = Num<W>::fms(st.p[(j + i) & (NP - 1)], st.co[i - 1], y);
      st.yh[j] = y;
    } else {
      // steps at or past T have g = 0 and leave the zero state as it is
      const W lam = (t < st.T ? v : Num<W>::zero()) + st.p[j];
      st.p[j] = Num<W>::zero();
#pragma unroll
      for (int i = 1; i <= NP; ++i) st.p[(j - i) & (NP - 1)] = Num<W>::fms(st.p[(j - i) & (NP - 1)], st.co[i - 1], lam);
    }
    pin_step();
    zsum_steps<S, NP, VEC, FWD, JJ + 1>(st, tb);
  }
}

template <typename S, int NP, bool FWD, bool VEC>
__global__ __launch_bounds__(256) void ti_zsum(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  using RC = RerunCfg<S, NP>;
  const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t cps = a.P - 1;
  if (item >= a.B * cps) return;
  const int64_t b = item / cps;
  const int64_t c = item - b * cps + (FWD ? 0 : 1);
  const int64_t t0 = c * a.L;
  ZsumLane<S, NP, VEC> st;
  load_coefs<S, NP>(st.co, a, b, !FWD);
#pragma unroll
  for (int r = 0; r < NP; ++r) {
    st.p[r] = Num<W>::zero();
    st.yh[r] = Num<W>::zero();
  }
  st.inp = a.x + b * a.T;
  st.T = a.T;
  const int64_t tfirst = FWD ? t0 : t0 + a.L - 1;
  if constexpr (VEC) {
#pragma unroll
    for (int q = 0; q < RC::DS / 4; ++q)
      st.xq[q] = opaque_v(ld4(st.inp, FWD ? min(tfirst + 4 * q, a.T - 4) : clamp64(tfirst - 3 - 4 * q, 0, a.T - 4)));
  } else {
#pragma unroll
    for (int k = 0; k < RC::DS; ++k)
      st.xs[k] = opaque_v(st.inp[FWD ? min(tfirst + k, a.T - 1) : clamp64(tfirst - k, 0, a.T - 1)]);
  }
  for (int jb = 0; jb < a.L; jb += RC::K) zsum_steps<S, NP, VEC, FWD, 0>(st, FWD ? t0 + jb : tfirst - jb);
  W* out = a.summ + item * NP;
  // the state at the chunk's far end: y_{t0+L-1-r} (forward), the pending sums of
  // positions t0-1-r (adjoint); both sit in slot NP-1-r
#pragma unroll
  for (int r = 0; r < NP; ++r) out[r] = FWD ? st.yh[NP - 1 - r] : st.p[NP - 1 - r];
}

template <typename S, int NP, bool FWD>
__global__ __launch_bounds__(64) void carry_ti(ChunkArgs<S> a) {
  using W = typename Wide<S>::type;
  __shared__ W phi[NP][NP + 1];
  __shared__ W sv[NP];
  const int64_t b = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t P = a.P;
  if (!FWD && a.ypad && lane < ti_ypad(NP)) {
    // the adjoint rerun's y window below position 0: element i is position
    // i - ti_ypad(NP), y_{-1-k} = zi[k] for k < n, zero beyond
    const int k = ti_ypad(NP) - 1 - lane;
    a.ypad[b * ti_ypad(NP) + lane] = (a.zi && k < a.n) ? a.zi[b * a.n + k] : Num<S>::zero();
  }
  if (lane < NP) {
    // lane k: response to y_{-1-k} = 1 over one whole chunk (column k of Phi)
    W co[NP];
    load_coefs<S, NP>(co, a, b, false);
    W ring[NP];
#pragma unroll
    for (int r = 0; r < NP; ++r) ring[r] = (r == NP - 1 - lane) ? Num<W>::one() : Num<W>::zero();
    for (int jb = 0; jb < a.L; jb += NP) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        W acc0 = Num<W>::zero(), acc1 = Num<W>::zero();
#pragma unroll
        for (int i = 1; i <= NP; i += 2) {
          acc0 = Num<W>::fms(acc0, co[i - 1], ring[(j - i) & (NP - 1)]);
          if (i < NP) acc1 = Num<W>::fms(acc1, co[i], ring[(j - i - 1) & (NP - 1)]);
        }
        ring[j] = acc0 + acc1;
      }
    }
#pragma unroll
    for (int r = 0; r < NP; ++r) phi[r][lane] = ring[NP - 1 - r];
  }
  __syncthreads();
  // Lanes NP..63 shadow lane NP-1 (same row, same values), so the chain's
  // stores need no branch: a store under a divergent branch makes the wait-count
  // pass drain the prefetch ring at every link.
  const int rl = lane < NP ? lane : NP - 1;
  W s = Num<W>::zero();
  if (FWD && rl < a.n && a.zi) s = convert<W>(a.zi[b * a.n + rl]);
  W* carry = a.carry + b * P * NP;
  const W* zs = a.summ + b * (P - 1) * NP;
  // Each lane keeps its row of the link matrix (Phi, or Phi^H for the adjoint)
  // in registers, and the links' zero-state vectors are prefetched ZD links
  // ahead into a register ring, by every lane at a clamped slot (forward link
  // `step` reads chunk slot step, adjoint link `step` chunk P-1-step at slot
  // P-2-step), so that no load sits under a branch. A slot is read after the
  // link's products and reloaded after that read, so old and new values never
  // overlap in its register; the barriers fence LDS only.
  W mrow[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) mrow[k] = FWD ? phi[rl][k] : Num<W>::conj(phi[k][rl]);
  constexpr int ZD = 8;
  auto zaddr = [&](int64_t step) -> const W* {
    const int64_t q = clamp64(FWD ? step : P - 2 - step, 0, P - 2);
    return zs + q * NP + rl;
  };
  W zr[ZD];
  if (P > 1) {
#pragma unroll
    for (int d = 0; d < ZD; ++d) zr[d] = opaque_v(*zaddr(d));
  }
  for (int64_t sb = 0; sb < P; sb += ZD) {
#pragma unroll
    for (int d = 0; d < ZD; ++d) {
      const int64_t step = sb + d;
      if (step >= P) break;
      const int64_t c = FWD ? step : P - 1 - step;
      carry[c * NP + rl] = s;
      sv[rl] = s;
      if (step == P - 1) break;
      lds_sync();
      W acc0 = Num<W>::zero(), acc1 = Num<W>::zero();
#pragma unroll
      for (int k = 0; k < NP; k += 2) {
        acc0 = Num<W>::fma(acc0, mrow[k], sv[k]);
        if (k + 1 < NP) acc1 = Num<W>::fma(acc1, mrow[k + 1], sv[k + 1]);
      }
      s = (acc0 + acc1) + zr[d];
      zr[d] = *zaddr(step + ZD);
      lds_sync();
    }
  }
}

// ----------------------------------------------------------------------------
// Host-side launch of the three phases.
// Summary work items (see fill_items) and chunk count of a batch.
template <typename S>
int64_t summary_items(const ChunkArgs<S>& a) {
  return a.rg ? a.G : a.B * (a.P - 1);
}
template <typename S>
int64_t chunk_count(const ChunkArgs<S>& a) {
  return a.rg ? a.G : a.B * a.P;
}

// float4 streams of the time-invariant kernels: float32 data, whole 4-sample
// blocks (T a multiple of 4; chunk lengths are multiples of 16) and 16-byte
// aligned streams (a null output counts as aligned).
template <typename S>
constexpr bool kF32 = std::is_same<S, float>::value;  // the float4 kernels exist for float32 only
template <typename S>
bool ti_vec(const ChunkArgs<S>& a, const void* in, const void* out) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if constexpr (std::is_same<S, float>::value)
    return a.T % 4 == 0 && a.L % 4 == 0 && al(in) && (out == nullptr || al(out));
  return false;
}

// The float4-stream reruns of time-varying float32 filters of order 2, 4, 8:
// uniform batch, n == NP, contiguous rows, T a multiple of 4, 16-byte aligned
// streams (x or g, y, A, and the outputs that are present).
template <typename S, int NP>
constexpr bool kTv4 = std::is_same<S, float>::value && (NP == 2 || NP == 4 || NP == 8);
// The prefetching multi-lane reruns (tlpc_grp_rerun.hip for orders 8, 16 and 32,
// tlpc_quad_rerun.hip for orders 2 and 4) exist for float32.
template <typename S, int NP>
constexpr bool kGrp = std::is_same<S, float>::value && (NP == 2 || NP == 4 || NP == 8 || NP == 16 || NP == 32);
// The multi-lane summaries exist for float32 orders 16 and 32 (DPP row
// broadcasts, tlpc_row_summary.hip) and 2, 4 and 8 (quads, tlpc_quad_rerun.hip).
template <typename S, int NP>
constexpr bool kRowSum = std::is_same<S, float>::value && (NP == 2 || NP == 4 || NP == 8 || NP == 16 || NP == 32);
template <typename S>
bool tv4_ok(const ChunkArgs<S>& a, int np, const void* y, const void* dx, const void* dA) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  return !a.rg && !a.neg && a.n == np && a.a_tstride == a.n && a.dapart == nullptr && a.T % 4 == 0 && a.L % 16 == 0 &&
         al(a.x) && al(a.A) && al(y) && (!dx || al(dx)) && (!dA || al(dA));
}

// Workgroup size of the reruns (one lane per chunk). A launch of few chunks (an
// order group of a ragged batch, a short batch) would fill few compute units
// with 256-lane workgroups; under 64K chunks it runs 64-lane ones, which spread
// the same lanes over four times as many.
inline unsigned rerun_threads(int64_t nch) { return nch < 256 * 256 ? 64u : 256u; }

template <typename S, int NP>
int launch_chunked_fwd(ChunkArgs<S> a, hipStream_t st) {
  using C = ChunkCfg<S, NP>;
  if (a.a_tstride == 0 && !a.rg) {
    const int64_t nz = a.B * (a.P - 1);
    const dim3 gz((unsigned)((nz + 255) / 256));
    if (nz > 0) {
      if (ti_vec(a, a.x, nullptr))
        hipLaunchKernelGGL((ti_zsum<S, NP, true, kF32<S>>), gz, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((ti_zsum<S, NP, true, false>), gz, dim3(256), 0, st, a);
    }
    hipLaunchKernelGGL((carry_ti<S, NP, true>), dim3((unsigned)a.B), dim3(64), 0, st, a);
    const int64_t nch = a.B * a.P;
    const unsigned rt = rerun_threads(nch);
    const dim3 g((unsigned)((nch + rt - 1) / rt));
    if (ti_vec(a, a.x, a.y))
      hipLaunchKernelGGL((fwd_rerun_ti<S, NP, kF32<S>>), g, dim3(rt), 0, st, a);
    else
      hipLaunchKernelGGL((fwd_rerun_ti<S, NP, false>), g, dim3(rt), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
  }
  if (a.P > 1) {
    bool done = false;
    if constexpr (kRowSum<S, NP>) {
      if (row_summary_ok(a, NP)) {
        const int rc = launch_row_summary(a, NP, true, st);
        if (rc != 0) return rc;
        done = true;
      }
    }
    if (!done) {
      using CF = ChunkCfg<S, NP, NP>;
      const int64_t grid = (summary_items(a) + CF::G - 1) / CF::G;
      hipLaunchKernelGGL((fwd_summary<S, NP>), dim3((unsigned)grid), dim3(CF::THREADS), 0, st, a);
    }
  }
  bool blocked = false;
  if constexpr (blocked_fits<S, NP>()) {
    if (a.P > 64) {
      launch_carry_blocked<S, NP, true>(a, st);
      blocked = true;
    }
  }
  if (!blocked)
    hipLaunchKernelGGL((carry_pipe<S, NP, true>), dim3((unsigned)a.B), dim3(256), 0, st, a);
  const int64_t nch = chunk_count(a);
  const unsigned rt = rerun_threads(nch);
  const dim3 g((unsigned)((nch + rt - 1) / rt));
  if constexpr (kTv4<S, NP>) {
    if (tv4_ok(a, NP, a.y, nullptr, nullptr)) {
      hipLaunchKernelGGL((fwd_rerun_tv4<NP>), g, dim3(rt), 0, st, a);
      return hipGetLastError() == hipSuccess ? 0 : -4;
    }
  }
  if constexpr (kGrp<S, NP>) {
    if (grp_rerun_ok(a, NP)) return launch_grp_fwd_rerun(a, NP, st);
  }
  hipLaunchKernelGGL((fwd_rerun<S, NP>), g, dim3(rt), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <typename S, int NP>
int launch_chunked_bwd(ChunkArgs<S> a, hipStream_t st) {
  using C = ChunkCfg<S, NP>;
  if (a.a_tstride == 0 && !a.rg) {
    const int64_t nz = a.B * (a.P - 1);
    const dim3 gz((unsigned)((nz + 255) / 256));
    if (nz > 0) {
      if (ti_vec(a, a.x, nullptr))
        hipLaunchKernelGGL((ti_zsum<S, NP, false, kF32<S>>), gz, dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((ti_zsum<S, NP, false, false>), gz, dim3(256), 0, st, a);
    }
    hipLaunchKernelGGL((carry_ti<S, NP, false>), dim3((unsigned)a.B), dim3(64), 0, st, a);
    const int64_t nch = a.B * a.P;
    const unsigned rt = rerun_threads(nch);
    const dim3 g((unsigned)((nch + rt - 1) / rt));
    // with da partials the y window moves in float4 blocks too (y aligned, and
    // the zi blocks in the workspace)
    const bool vec = ti_vec(a, a.x, a.dx) &&
                     (!a.dapart || (a.ypad && (reinterpret_cast<uintptr_t>(a.y) & 15) == 0));
    // a full per-sample dA of a time-invariant filter takes the per-step row
    // kernel, which reads the one row at stride 0
    if (a.dapart) {
      if (vec)
        hipLaunchKernelGGL((bwd_rerun_ti<S, NP, DA_PARTIAL, kF32<S>>), g, dim3(rt), 0, st, a);
      else
        hipLaunchKernelGGL((bwd_rerun_ti<S, NP, DA_PARTIAL, false>), g, dim3(rt), 0, st, a);
    } else if (a.dA) {
      hipLaunchKernelGGL((bwd_rerun<S, NP>), g, dim3(rt), 0, st, a);
    } else {
      if (vec)
        hipLaunchKernelGGL((bwd_rerun_ti<S, NP, DA_NONE, kF32<S>>), g, dim3(rt), 0, st, a);
      else
        hipLaunchKernelGGL((bwd_rerun_ti<S, NP, DA_NONE, false>), g, dim3(rt), 0, st, a);
    }
    if (a.dapart) {
      const int64_t nd = a.B * a.n;
      hipLaunchKernelGGL((ti_da_reduce<S, NP>), dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
  }
  bool rowsum = false;
  bool kept = false;
  if constexpr (kRowSum<S, NP>) {
    // with the forward's summaries at hand only the zero-state responses are new
    if (a.P > 1 && a.kept && grp_wsum_ok(a, NP)) {
      const int rc = launch_grp_wsum(a, NP, st);
      if (rc != 0) return rc;
      rowsum = kept = true;
    } else if (a.P > 1 && row_summary_ok(a, NP)) {
      const int rc = launch_row_summary(a, NP, false, st);
      if (rc != 0) return rc;
      rowsum = true;
    }
  }
  a.kept = kept ? 1 : 0;  // the chain reads the slots transposed only when they hold Phi
  if (a.P > 1 && !rowsum) {
    if constexpr (NP == 32) {
      using CB = ChunkCfg<S, NP, NP>;
      const int64_t grid = (summary_items(a) + CB::G - 1) / CB::G;
      hipLaunchKernelGGL((bwd_summary_ff<S, NP>), dim3((unsigned)grid), dim3(CB::THREADS), 0, st, a);
    } else {
      const int64_t grid = (summary_items(a) + C::G - 1) / C::G;
      hipLaunchKernelGGL((bwd_summary<S, NP>), dim3((unsigned)grid), dim3(C::THREADS), 0, st, a);
    }
  }
  bool blocked = false;
  if constexpr (blocked_fits<S, NP>()) {
    if (a.P > 64) {
      launch_carry_blocked<S, NP, false>(a, st);
      blocked = true;
    }
  }
  if (!blocked)
    hipLaunchKernelGGL((carry_pipe<S, NP, false>), dim3((unsigned)a.B), dim3(256), 0, st, a);
  const int64_t nch = chunk_count(a);
  if (a.rx) {
    // y rebuilt from the forward's carries (real data, orders up to 4)
    if constexpr (NP <= 4 && (std::is_same<S, float>::value || std::is_same<S, double>::value)) {
      if (a.rg || a.dapart || a.L > RC_LMAX || a.L % RC_SB != 0) return -1;
      hipLaunchKernelGGL((bwd_rerun_rc<S, NP>), dim3((unsigned)((nch + 63) / 64)), dim3(64), 0, st, a);
      return hipGetLastError() == hipSuccess ? 0 : -4;
    } else {
      return -1;
    }
  }
  const unsigned rt = rerun_threads(nch);
  const dim3 g((unsigned)((nch + rt - 1) / rt));
  bool tv4 = false;
  if constexpr (kTv4<S, NP>) {
    if (tv4_ok(a, NP, a.y, a.dx, a.dA)) {
      hipLaunchKernelGGL((bwd_rerun_tv4<NP>), g, dim3(rt), 0, st, a);
      tv4 = true;
    }
  }
  bool grp = false;
  if constexpr (kGrp<S, NP>) {
    if (!tv4 && grp_rerun_ok(a, NP)) {
      const int rc = launch_grp_bwd_rerun(a, NP, st);
      if (rc != 0) return rc;
      grp = true;
    }
  }
  if (!tv4 && !grp) hipLaunchKernelGGL((bwd_rerun<S, NP>), g, dim3(rt), 0, st, a);
  if (a.dapart) {
    const int64_t nd = a.B * a.n;
    hipLaunchKernelGGL((ti_da_reduce<S, NP>), dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

#define TLPC_INST(S, NP)                                             \
  template int launch_chunked_fwd<S, NP>(ChunkArgs<S>, hipStream_t); \
  template int launch_chunked_bwd<S, NP>(ChunkArgs<S>, hipStream_t);

#define TLPC_INST_ALL_NP(S) \
  TLPC_INST(S, 1)           \
  TLPC_INST(S, 2)           \
  TLPC_INST(S, 4)           \
  TLPC_INST(S, 8)           \
  TLPC_INST(S, 16)          \
  TLPC_INST(S, 32)

}  // namespace tlpc
