// Per-signal tiled filter for order 16, float32 (TLPC_STRATEGY_SIGNAL): forward
// and fused adjoint, each reading every sample of A, x / g and y from HBM once.
//
// One workgroup of 8 waves owns one signal and walks it in tiles (the backward
// from the last tile down) of sub-chunks of LC = 32 samples: 12 sub-chunks (384
// samples) per tile in the forward, 16 (512 samples) in the backward. Each tile
// passes through three stages (four in the backward), one per period of a
// software pipeline, so that in period p the waves work on several tiles at once:
//
//   stage 1 (tile p)    one 16-lane DPP row per sub-chunk (forward waves 0-2,
//                       backward waves 0-3) sweeps it backwards with the
//                       transposed recurrence from a unit at output r (lane r):
//                       row r of the sub-chunk's transition matrix Phi, and in
//                       the forward z[r] (the lambdas met on the way, dotted with
//                       x), in float64.
//                       backward, wave 4: one quad per sub-chunk runs the adjoint
//                       recurrence from a zero state through it (its zero-state
//                       response w).
//   stage 2 (tile p-1)  one wave: the carry chain over the tile's sub-chunks,
//                       continuing the previous tile's state in registers:
//                       s <- Phi s + z (forward), s <- Phi^T s + w (backward,
//                       descending).
//   stage 3 (tile p-2)  one wave: one quad per sub-chunk reruns it from its carry
//                       in float64 arithmetic on the float32 data and writes y
//                       (forward) or lambda into LDS (backward).
//   stage 4 (tile p-3)  backward, one wave: dx, the 64-byte dA rows and, at the
//                       start of the signal, dzi.
//
// A tile's A rows and per-sample streams sit in LDS from stage 1 to stage 3;
// the next tile's global loads are issued into registers when a period starts
// and written to LDS later, so they overlap the period's arithmetic. The
// backward (three row buffers, one summary buffer) writes them, and stage 1 its
// summaries, after a barrier at the end of the period, when stage 2 is done
// with the previous tile's summaries. The forward (four row buffers, two summary
// buffers) has each wave write them as soon as its own stage is done, and ends
// a period at one barrier.
//
// Samples past the end of the signal are zero rows and zero inputs in LDS, so
// every sub-chunk is whole: the forward computes values there that it does not
// store; in the adjoint they are steps with lambda = 0.
//
// Summaries and carries never leave the chip, so the strategy needs no
// workspace. Reference recurrence: torchlpc/csrc/scan_cpu.cpp:112-120;
// adjoint: torchlpc/core.py:186-229.
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"

namespace tlpc {
namespace t16 {

using namespace dpp16;

constexpr int NP = 16;
constexpr int S = 512;          // samples per tile
constexpr int LC = 32;          // samples per sub-chunk
constexpr int C = S / LC;       // sub-chunks per tile (one DPP row / quad each)
constexpr int THREADS = 512;    // 8 waves
constexpr int NBUF = 3;         // tiles resident in LDS (one per pipeline stage)
constexpr int ROWP = NP + 1;    // Phi rows padded to 17 doubles in LDS
constexpr int LINKP = NP * ROWP + NP;
// Every thread stages SA float4 of the next tile's rows (the first SX threads
// also its per-sample streams), every period, with every load unconditional at
// a clamped index: a load under any condition leaves the compiler a merge of the
// loaded registers, which it resolves with copies that wait for the load right
// there. (The waves that store also wait for their stores at the commit: loads
// and stores share one counter on this GPU.)
constexpr int STAGERS = THREADS;
constexpr int SX = S / 4;                 // float4 of a per-sample stream per tile
constexpr int SA = S * NP / 4 / STAGERS;  // float4 of A per staging thread
static_assert(C == 16 && LC % 16 == 0 && SA * STAGERS * 4 == S * NP && SX <= STAGERS, "tile layout");
// The forward's tiles hold 12 sub-chunks (384 samples): three waves sweep Phi,
// one per SIMD, and the fourth SIMD runs the quad rerun and the carry chain,
// whose serial steps then do not wait behind a Phi wave's issue (the Phi sweep
// keeps its SIMD's float64 pipe busy). With 16 sub-chunks, four Phi waves and
// the rerun sharing a SIMD with one of them, a 512-sample period took about
// 1.65 times a 384-sample one.
constexpr int FS = 384;
constexpr int FC = FS / LC;
constexpr int FSX = FS / 4;
static_assert(FC % 4 == 0 && FC <= 16 && (FS * NP / 4) % STAGERS == 0 && FSX <= STAGERS, "forward tile layout");

// The forward keeps four tiles of rows in LDS and two of summaries, so that no
// buffer a wave writes in a period is read in it: tile p+1's rows go where tile
// p-3's were (its rerun ended the period before), tile p's summaries where
// tile p-2's were (the chain read them the period before). Each wave then
// commits its staged loads, and the sweeping waves their summaries, as soon as
// its own stage is done, and a period ends at a single barrier (156 KB of LDS).
constexpr int FNBUF = 4;
struct FwdLds {
  float4 A[FNBUF][FS * 4];    // tile rows: A[b][r*4 + q] = A[t0 + r][4q .. 4q+3]
  float4 x[FNBUF][FSX];       // x[t0 .. t0+FS)
  double summ[2][FC][LINKP];  // Phi rows (padded) then z: tile p's in summ[p & 1]
  double carry[2][FC][NP];    // state entering each sub-chunk (written in stage 2, read in 3)
};

struct BwdLds {
  float4 A[NBUF][S * 4];
  float4 g[NBUF][SX];         // g[t0 .. t0+S)
  float4 y[NBUF + 1][SX + 4]; // y[t0-16 .. t0+S) (y[-1-k] = zi[k] below the signal), stages 1-4
  double summ[C][LINKP];      // Phi rows (padded) then w
  double carry[2][C][NP];     // adjoint state entering each sub-chunk from above
  float lam[2][S];            // lambda of a tile, from stage 3 to stage 4
};

__device__ __forceinline__ void lds_barrier() {
  // orders LDS only: the staged global loads stay in flight across it
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ float4 zero4() { return make_float4(0.f, 0.f, 0.f, 0.f); }

// float4 of floats [i, i+4) of a stream of length n (n >= 1): one 16-byte load
// when the stream is 16-byte aligned (ALIGNED), else four loads, always at
// clamped indices. Elements at or past n are zeroed by zfix where the value is
// used, not here: writing the loaded registers under a condition makes the
// compiler wait for the load on the spot.
template <bool ALIGNED>
__device__ __forceinline__ float4 ld4(const float* p, int64_t i, int64_t n) {
  if constexpr (ALIGNED) {
    return *reinterpret_cast<const float4*>(p + (i < n ? i : n - 4));
  } else {
    float4 v;
    v.x = p[i + 0 < n ? i + 0 : n - 1];
    v.y = p[i + 1 < n ? i + 1 : n - 1];
    v.z = p[i + 2 < n ? i + 2 : n - 1];
    v.w = p[i + 3 < n ? i + 3 : n - 1];
    return v;
  }
}
__device__ __forceinline__ float4 zfix(float4 v, int64_t i, int64_t n) {
  v.x = i + 0 < n ? v.x : 0.f;
  v.y = i + 1 < n ? v.y : 0.f;
  v.z = i + 2 < n ? v.z : 0.f;
  v.w = i + 3 < n ? v.w : 0.f;
  return v;
}
template <bool ALIGNED>
__device__ __forceinline__ float4 ld4z(const float* p, int64_t i, int64_t n) {
  return zfix(ld4<ALIGNED>(p, i, n), i, n);
}

// Stores floats [i, i+4) of a stream of length n (those below n).
template <bool ALIGNED>
__device__ __forceinline__ void st4(float* p, int64_t i, int64_t n, float4 v) {
  if (ALIGNED && i + 4 <= n) {
    *reinterpret_cast<float4*>(p + i) = v;
  } else {
    if (i + 0 < n) p[i + 0] = v.x;
    if (i + 1 < n) p[i + 1] = v.y;
    if (i + 2 < n) p[i + 2] = v.z;
    if (i + 3 < n) p[i + 3] = v.w;
  }
}

// Staging of one tile: staging thread tid (< STAGERS) holds float4 tid + STAGERS*j
// of the tile's A rows (rows past T are zero) and float4 `tid` of one per-sample
// stream.
// Makes the wave wait for a staged load here, on every path: the wait-count pass
// then sees no load pending into these registers at the loop back edge (it
// would otherwise wait for every load in flight before the next period's
// staging loads reuse them).
__device__ __forceinline__ void consume(float4& v) { asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w)); }

template <bool ALIGNED, int TS = S>
struct Stage {
  static constexpr int TSA = TS * NP / 4 / STAGERS;  // float4 of A per staging thread
  static constexpr int TSX = TS / 4;                 // float4 of a per-sample stream
  float4 a[TSA];
  float4 s0;      // x (forward) or g (backward): float4 tid % TSX of the tile
  float4 sy, sh;  // backward: y (float4 tid % TSX) and the y halo (float4 tid % 4)
  int64_t t0;     // first sample of the staged tile
  // Waits for every staged load; called on every path once per period.
  __device__ __forceinline__ void consume_all() {
#pragma unroll
    for (int j = 0; j < TSA; ++j) consume(a[j]);
    consume(s0);
    consume(sy);
    consume(sh);
  }
  // Staging thread tid in [0, STAGERS).
  __device__ __forceinline__ void load_rows(const float* Ab, int64_t T, int64_t t0_, int tid) {
    t0 = t0_;
    const int64_t n = T * NP;
#pragma unroll
    for (int j = 0; j < TSA; ++j) a[j] = ld4<ALIGNED>(Ab, t0 * NP + 4 * (int64_t)(tid + STAGERS * j), n);
  }
  // Tiles wholly inside the signal (every one but the last) store the staged
  // values as they are; the last one zeroes what lies past the end.
  __device__ __forceinline__ bool whole(int64_t T) const { return t0 + TS <= T; }
  __device__ __forceinline__ void commit_rows(float4* As, int64_t T, int tid) {
    const int64_t n = T * NP;
    if (whole(T)) {
#pragma unroll
      for (int j = 0; j < TSA; ++j) As[tid + STAGERS * j] = a[j];
    } else {
#pragma unroll
      for (int j = 0; j < TSA; ++j) As[tid + STAGERS * j] = zfix(a[j], t0 * NP + 4 * (int64_t)(tid + STAGERS * j), n);
    }
  }
  // Forward x of the staged tile (threads past SX load a duplicate).
  __device__ __forceinline__ void load_x(const float* xb, int64_t T, int tid) {
    s0 = ld4<ALIGNED>(xb, t0 + 4 * (tid % TSX), T);
  }
  __device__ __forceinline__ void commit_x(float4* xs, int64_t T, int tid) {
    if (tid < TSX) xs[tid] = whole(T) ? s0 : zfix(s0, t0 + 4 * tid, T);
  }
  // Backward streams of the staged tile: g, y and the 16-sample y halo below the
  // tile (zi reversed below the signal start: y_{-16+4j+m} = zi[15-4j-m]).
  __device__ __forceinline__ void load_gy(const float* gb, const float* yb, const float* zib, int64_t T,
                                          int tid) {
    s0 = ld4<ALIGNED>(gb, t0 + 4 * (tid % TSX), T);
    sy = ld4<ALIGNED>(yb, t0 + 4 * (tid % TSX), T);
    // the halo: y below the tile, or zi (or g again, never used, without zi)
    const float* hp = t0 > 0 ? yb : (zib ? zib : gb);
    const int64_t hi = t0 > 0 ? t0 - 16 + 4 * (tid % 4) : 12 - 4 * (tid % 4);
    sh = ld4<false>(hp, hi, t0 > 0 ? T : NP);
  }
  __device__ __forceinline__ void commit_gy(float4* gs, float4* ys, const float* zib, int64_t T, int tid) {
    consume(s0);
    consume(sy);
    consume(sh);
    if (tid < TSX) {
      if (whole(T)) {
        gs[tid] = s0;
        ys[4 + tid] = sy;
      } else {
        gs[tid] = zfix(s0, t0 + 4 * tid, T);
        ys[4 + tid] = zfix(sy, t0 + 4 * tid, T);
      }
    }
    if (tid < 4) {
      const float4 v = zib ? sh : zero4();
      ys[tid] = t0 > 0 ? sh : make_float4(v.w, v.z, v.y, v.x);
    }
  }
};

// ---------------------------------------------------------------------------
// Phase 1: transition-matrix rows of one sub-chunk per 16-lane DPP row.
//
// Lane kk starts from a unit at output position t0c+LC-1-kk and sweeps the
// sub-chunk backwards with the transposed recurrence (adj_row_step4): the
// pending sums left for the 16 positions before the sub-chunk are row kk of Phi,
// and with WITH_Z the lambdas met, dotted with x, give z[kk]. Steps 4M+3 .. 4M of
// a 16-step block use the float4 group M: lane kk holds row (block + 4M + kk/4),
// coefficients 4(kk%4) .. 4(kk%4)+3. A group is converted, then reloaded (from
// LDS) for the block below, so the LDS read overlaps four steps of arithmetic.
template <int J, bool WITH_Z>
__device__ __forceinline__ void row_quad(double (&p)[NP], double& z, const double (&d)[4], double xv) {
  if constexpr (J >= 0) {
    const double lam = adj_row_step4<J>(p, d);
    if constexpr (WITH_Z) fma_bcast<J>(z, xv, lam);
    if constexpr ((J & 3) != 0) row_quad<J - 1, WITH_Z>(p, z, d, xv);
  }
}
template <int M, bool WITH_Z>
__device__ __forceinline__ void row_block(double (&p)[NP], double& z, float4 (&cq)[4], double xv,
                                          const float4* nq) {
  if constexpr (M >= 0) {
    const double d[4] = {(double)cq[M].x, (double)cq[M].y, (double)cq[M].z, (double)cq[M].w};
    cq[M] = nq[M * 16];
    row_quad<4 * M + 3, WITH_Z>(p, z, d, xv);
    row_block<M - 1, WITH_Z>(p, z, cq, xv, nq);
  }
}

// Sub-chunk c of the tile in LDS (rows As, samples xs): Phi rows and z into
// registers (p[NP - 1 - k] = Phi[kk][k]).
template <bool WITH_Z>
__device__ __forceinline__ void phi_sweep(const float4* As, const float* xs, int c, int kk,
                                          double (&p)[NP], double& z) {
  const float4* ap = As + c * LC * 4 + kk;
#pragma unroll
  for (int i = 0; i < NP; ++i) p[i] = (i == NP - 1 - kk) ? 1.0 : 0.0;
  z = 0.0;
  float4 cq[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) cq[m] = ap[(LC - NP + 4 * m) * 4];
  float xf = WITH_Z ? xs[c * LC + LC - NP + kk] : 0.f;
#pragma unroll 1
  for (int jb = LC - NP; jb >= 0; jb -= NP) {
    const int jn = jb >= NP ? jb - NP : 0;  // the block below (block 0 again after the last)
    const double xv = (double)xf;
    float xn = 0.f;
    if constexpr (WITH_Z) xn = xs[c * LC + jn + kk];
    row_block<3, WITH_Z>(p, z, cq, xv, ap + jn * 4);
    xf = xn;
  }
}

template <bool WITH_Z>
__device__ __forceinline__ void phi_store(double* out, int kk, const double (&p)[NP], double z) {
#pragma unroll
  for (int k = 0; k < NP; ++k) out[kk * ROWP + k] = p[NP - 1 - k];  // Phi[kk][k]
  if constexpr (WITH_Z) out[NP * ROWP + kk] = z;
}

// ---------------------------------------------------------------------------
// Phase 2: the chain over the C sub-chunks of a tile on lanes 0-15 (the other
// rows of the wave repeat row 0). Forward: carry[c] = s; s = Phi_c s + z_c for
// c = 0..C-1. Backward: carry[c] = s; s = Phi_c^T s + w_c for c = C-1..0. The
// LDS reads of the next link go out before the current link's products.
template <bool FWD, int CN = C>
__device__ __forceinline__ void chain(double& s, const double (*summ)[LINKP], double (*carry)[NP], int lane) {
  const int row = lane & 15;
  double m[2][NP], zv[2];
  auto ldm = [&](int c, double(&mm)[NP], double& z) {
#pragma unroll
    for (int k = 0; k < NP; ++k) mm[k] = FWD ? summ[c][row * ROWP + k] : summ[c][k * ROWP + row];
    z = summ[c][NP * ROWP + row];
  };
  ldm(FWD ? 0 : CN - 1, m[0], zv[0]);
#pragma unroll
  for (int i = 0; i < CN; ++i) {
    const int c = FWD ? i : CN - 1 - i;
    if (i + 1 < CN) ldm(FWD ? c + 1 : c - 1, m[(i + 1) & 1], zv[(i + 1) & 1]);
    if (lane < NP) carry[c][lane] = s;
    s = chain_link(s, m[i & 1], zv[i & 1]);
  }
}

// ---------------------------------------------------------------------------
// Stage 3: quad q reruns sub-chunk q (LC steps) from its carry in float64
// arithmetic on the float32 data (a float32 rerun, even over 32 steps from a
// float64 carry, left y at 1e-5 of FP64 on the B=256, T=48000, N=16 inputs).
constexpr int qbcast_ctrl(int q) { return q * 0x55; }  // quad_perm [q,q,q,q]
template <int K>
__device__ __forceinline__ float comp(const float4& v) {
  if constexpr ((K & 3) == 0) return v.x;
  if constexpr ((K & 3) == 1) return v.y;
  if constexpr ((K & 3) == 2) return v.z;
  return v.w;
}
template <int K>
__device__ __forceinline__ float qsel(const float4& v) {  // sample K of a block, in every lane
  return qperm<qbcast_ctrl(K >> 2)>(comp<K>(v));
}
template <int K>
__device__ __forceinline__ void put(float4& o, int ql, float v) {  // sample K of a block, owner lane
  const bool mine = ql == (K >> 2);
  if constexpr ((K & 3) == 0) o.x = mine ? v : o.x;
  if constexpr ((K & 3) == 1) o.y = mine ? v : o.y;
  if constexpr ((K & 3) == 2) o.z = mine ? v : o.z;
  if constexpr ((K & 3) == 3) o.w = mine ? v : o.w;
}

constexpr int RING = 4;  // A rows in flight per quad lane (LDS reads ahead of their step)
// Keeps a ring refill where the step issues it: without it the scheduler sank
// the LDS reads to a step or two before their use (and lent the ring's
// registers to temporaries), so the steps waited for their rows.
__device__ __forceinline__ void ring_pin() { __builtin_amdgcn_sched_barrier(0); }

// Forward quad step at time t (slot J = t mod 4): lane ql's window R holds y_u
// for u in [t-4-4ql, t-1-4ql] at slot u mod 4, c = A[t, 4ql .. 4ql+3]. The
// product with the newest value (slot (J-1) & 3: y_{t-1} in lane 0) goes last,
// so the step waits on it for one FMA. Returns y_t in every lane of the quad.
template <int J>
__device__ __forceinline__ double fwd_step(double (&R)[4], const float4 c, float x, int ql) {
  const double cc[4] = {(double)c.x, (double)c.y, (double)c.z, (double)c.w};
  constexpr int NW = (J - 1) & 3;
  constexpr int S1 = (NW + 1) & 3, S2 = (NW + 2) & 3, S3 = (NW + 3) & 3;
  double p = cc[(J - 1 - S1) & 3] * R[S1];
  p = fma(cc[(J - 1 - S2) & 3], R[S2], p);
  p = fma(cc[(J - 1 - S3) & 3], R[S3], p);
  p = fma(cc[(J - 1 - NW) & 3], R[NW], p);
  const double y = (double)x - quad_sum(p);
  const double up = qperm<QP_UP>(R[J]);  // lane ql-1 holds y_{t-4ql} in slot J
  R[J] = (ql == 0) ? y : up;
  return y;
}

template <int JJ, int TS>
__device__ __forceinline__ void fwd_steps(double (&R)[4], float4 (&cq)[RING], float4& out,
                                          const float4& xc, const float4* ar, int tb, int ql) {
  if constexpr (JJ < 16) {
    const double y = fwd_step<JJ & 3>(R, cq[JJ % RING], qsel<JJ>(xc), ql);
    cq[JJ % RING] = ar[min(tb + JJ + RING, TS - 1) * 4];
    ring_pin();
    put<JJ>(out, ql, (float)y);
    fwd_steps<JJ + 1, TS>(R, cq, out, xc, ar, tb, ql);
  }
}

// A tile of TS samples holds TS / LC sub-chunks; quads past them (TS < 16 LC)
// rerun the last sub-chunk again and store nothing.
template <bool ALIGNED, int TS>
__device__ __forceinline__ void fwd_rerun(const float4* As, const float4* xs4, const double (*carry)[NP],
                                          float* yb, int64_t T, int64_t t0, int lane) {
  constexpr int CN = TS / LC;
  const int qa = lane >> 2, ql = lane & 3;
  const int q = qa < CN ? qa : CN - 1;
  double R[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) R[3 - m] = carry[q][4 * ql + m];  // y_{t0c-1-4ql-m}, slot 3-m
  const float4* ar = As + ql;
  float4 cq[RING];
#pragma unroll
  for (int j = 0; j < RING; ++j) cq[j] = ar[(q * LC + j) * 4];
  float4 xc = xs4[(q * LC) / 4 + ql];
#pragma unroll 1
  for (int jb = 0; jb < LC; jb += 16) {
    const int tb = q * LC + jb;  // row in the tile
    const float4 xn = xs4[min(tb + 16, TS - 16) / 4 + ql];
    float4 out = zero4();
    fwd_steps<0, TS>(R, cq, out, xc, ar, tb, ql);
    if (qa < CN) st4<ALIGNED>(yb, t0 + tb + 4 * ql, T, out);
    xc = xn;
  }
}

// ---------------------------------------------------------------------------
// Backward quads over sub-chunk q, descending. Stage 1 (ZS_WAVE): the
// zero-state response w in float64 (it feeds the float64 chain). Stage 3
// (RERUN_WAVE): lambda from the sub-chunk's carry, into LDS (LamT: the type of
// its recurrence).
// Samples past the end of the signal are zero rows with zero g and zero state
// from above, so their lambda is 0 without a test.

// One tile of the backward in LDS: rows, g, and the y window y[t0-16 .. t0+S)
// (window index = tile row + 16).
struct BwdTile {
  const float4* A;
  const float4* g;
  const float4* y;
};

template <int JJ>
__device__ __forceinline__ void zs_steps(double (&P)[4], double& now, float4 (&cq)[RING], const float4& gc,
                                         const float4* ar, int tb, int ql) {
  if constexpr (JJ < 16) {
    constexpr int J = (3 - JJ) & 3;  // slot of row tb + 15 - JJ (tb is a multiple of 16)
    const int r = tb + 15 - JJ;
    adj_quad_step<J>(P, now, cq[JJ % RING], qsel<15 - JJ>(gc), ql, true);
    cq[JJ % RING] = ar[max(r - RING, 0) * 4];
    ring_pin();
    zs_steps<JJ + 1>(P, now, cq, gc, ar, tb, ql);
  }
}

__device__ __forceinline__ void zero_state(const BwdTile& L, double (&P)[4], double& now, int q, int ql) {
  const float4* ar = L.A + ql;
  const int top = q * LC + LC - 1;
  float4 cq[RING];
#pragma unroll
  for (int j = 0; j < RING; ++j) cq[j] = ar[(top - j) * 4];
#pragma unroll 1
  for (int jb = LC - 16; jb >= 0; jb -= 16) {
    const int tb = q * LC + jb;
    float4 gc = L.g[tb / 4 + ql];
    consume(gc);  // in registers for the block, not read from LDS again at every step
    zs_steps<0>(P, now, cq, gc, ar, tb, ql);
  }
}

// Adjoint quad step at time t (slot J = t mod 4), descending, in float32: `now`
// in lane 0 is the pending sum of position t; lane ql's window P holds the
// pending sums of positions [t-4-4ql, t-1-4ql] at slot u mod 4. Returns lambda_t
// (every lane) and leaves the state for step t-1.
template <int J>
__device__ __forceinline__ float adj_step(float (&P)[4], float& now, const float4 c, float g, int ql) {
  const float lam = qperm<QP_LANE0>(g + now);
  P[(J - 1) & 3] = fmaf(-c.x, lam, P[(J - 1) & 3]);
  P[(J - 2) & 3] = fmaf(-c.y, lam, P[(J - 2) & 3]);
  P[(J - 3) & 3] = fmaf(-c.z, lam, P[(J - 3) & 3]);
  P[(J - 4) & 3] = fmaf(-c.w, lam, P[(J - 4) & 3]);
  constexpr int SL = (J - 1) & 3;  // slot that leaves at the top and enters at the bottom
  const float down = qperm<QP_DOWN>(P[SL]);
  now = P[SL];
  P[SL] = (ql == 3) ? 0.f : down;
  return lam;
}
// the same in float64 (tlpc_dpp16.h)
template <int J>
__device__ __forceinline__ double adj_step(double (&P)[4], double& now, const float4 c, float g, int ql) {
  return adj_quad_step<J>(P, now, c, g, ql, true);
}

template <int JJ, typename V>
__device__ __forceinline__ void lam_steps(V (&P)[4], V& now, float4 (&cq)[RING], float4& lo,
                                          const float4& gc, const float4* ar, int tb, int ql) {
  if constexpr (JJ < 16) {
    constexpr int J = (3 - JJ) & 3;
    const int r = tb + 15 - JJ;
    const V lam = adj_step<J>(P, now, cq[JJ % RING], qsel<15 - JJ>(gc), ql);
    cq[JJ % RING] = ar[max(r - RING, 0) * 4];
    ring_pin();
    put<15 - JJ>(lo, ql, (float)lam);
    lam_steps<JJ + 1>(P, now, cq, lo, gc, ar, tb, ql);
  }
}

// lambda of sub-chunk q into lam[row] (LDS), from the state in P / now.
template <typename V>
__device__ __forceinline__ void lam_rerun(const BwdTile& L, float4* lam4, V (&P)[4], V& now, int q, int ql) {
  const float4* ar = L.A + ql;
  const int top = q * LC + LC - 1;
  float4 cq[RING];
#pragma unroll
  for (int j = 0; j < RING; ++j) cq[j] = ar[(top - j) * 4];
#pragma unroll 1
  for (int jb = LC - 16; jb >= 0; jb -= 16) {
    const int tb = q * LC + jb;
    float4 lo = zero4();
    float4 gc = L.g[tb / 4 + ql];
    consume(gc);  // in registers for the block, not read from LDS again at every step
    lam_steps<0>(P, now, cq, lo, gc, ar, tb, ql);
    lam4[tb / 4 + ql] = lo;
  }
}

// Stage 4 (DA_WAVE): dx = lambda and dA[t][i-1] = -lambda_t * y_{t-i} of a tile
// from LDS. Lane l takes rows l, l+64, ...: each store instruction of the wave
// writes 64 consecutive 64-byte dA rows.
template <bool ALIGNED>
__device__ __forceinline__ void grads(const float* lam, const float* ywin, float* dxb, float* dAb, int64_t T,
                                      int64_t t0, int lane) {
  if (dAb) {
    // all of a lane's rows in flight at once (their LDS reads issue together):
    // against two at a time, 1.5 % off the backward
#pragma unroll 8
    for (int j = 0; j < S / 64; ++j) {
      const int r = lane + 64 * j;
      const int64_t t = t0 + r;
      // every LDS read of the row first, then the products: the products
      // interleaved with the reads made the wave wait for LDS once per pair
      float yv[NP];
#pragma unroll
      for (int i = 0; i < NP; ++i) yv[i] = ywin[r + 15 - i];  // y_{t-1-i} at window r+15-i
      const float nl = -lam[r];
      asm volatile("" : "+v"(yv[0]), "+v"(yv[1]), "+v"(yv[2]), "+v"(yv[3]), "+v"(yv[4]), "+v"(yv[5]),
                   "+v"(yv[6]), "+v"(yv[7]), "+v"(yv[8]), "+v"(yv[9]), "+v"(yv[10]), "+v"(yv[11]),
                   "+v"(yv[12]), "+v"(yv[13]), "+v"(yv[14]), "+v"(yv[15]));
      float d[NP];
#pragma unroll
      for (int i = 0; i < NP; ++i) d[i] = nl * yv[i];
      if (t < T) {
        float* row = dAb + t * NP;
        if constexpr (ALIGNED) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            reinterpret_cast<float4*>(row)[v] = make_float4(d[4 * v], d[4 * v + 1], d[4 * v + 2], d[4 * v + 3]);
        } else {
#pragma unroll
          for (int i = 0; i < NP; ++i) row[i] = d[i];
        }
      }
    }
  }
  if (dxb) {
    const float4* l4 = reinterpret_cast<const float4*>(lam);
#pragma unroll
    for (int j = 0; j < SX / 64; ++j) st4<ALIGNED>(dxb, t0 + 4 * (lane + 64 * j), T, l4[lane + 64 * j]);
  }
}

// ---------------------------------------------------------------------------
// Roles in period p: the Phi waves stage 1 (tile p), the chain wave stage 2
// (tile p-1), the rerun wave stage 3 (tile p-2), and in the backward wave
// DA_WAVE stage 4 (tile p-3); every wave stages tile p+1. Wave w runs on SIMD
// w % 4: the forward's three Phi waves take SIMDs 0-2 and its rerun and chain
// SIMD 3; the backward's four Phi waves share SIMDs 0 and 1, its zero-state and
// lambda quads SIMD 2, its chain and dA stage SIMD 3. Tile p
// lives in buffer p % NBUF (the backward's y window in p % (NBUF + 1)), its
// carries in carry[p % 2].
// backward: Phi sweeps (stage 1) on waves 0, 1, 4, 5, two on each of SIMDs 0
// and 1; the serial stages take SIMDs 2 and 3
constexpr bool bwd_phi(int w) { return (w & 2) == 0; }
constexpr int bwd_phi_index(int w) { return (w & 1) | ((w >> 2) << 1); }
constexpr int ZS_WAVE = 2;       // backward: zero-state responses (stage 1)
constexpr int CHAIN_WAVE = 3;
constexpr int FWD_PHI_WAVES = FC / 4;  // forward: Phi sweeps (stage 1), waves 0-2
constexpr int FWD_RERUN_WAVE = 3;      // forward: rerun and y stores (stage 3), SIMD 3
constexpr int FWD_CHAIN_WAVE = 7;      // forward: carry chain (stage 2), SIMD 3
constexpr int RERUN_WAVE = 6;    // backward: lambda (stage 3)
constexpr int DA_WAVE = 7;       // backward: dA and dx stores (stage 4)


template <bool ALIGNED>
__global__ __launch_bounds__(THREADS) void tile_fwd(ChunkArgs<float> a) {
  __shared__ FwdLds L;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t b = blockIdx.x;
  const int64_t T = a.T;
  const int64_t ntiles = (T + FS - 1) / FS;
  const float* Ab = a.A + b * T * NP;
  const float* xb = a.x + b * T;
  float* yb = a.y + b * T;
  double s = 0.0;  // chain state (lanes 0-15 of the chain wave): y_{t0-1-r}
  if (wave == FWD_CHAIN_WAVE && lane < NP && a.zi) s = (double)a.zi[b * NP + lane];
  const int si = tid;
  {
    Stage<ALIGNED, FS> st;
    st.load_rows(Ab, T, 0, si);
    st.load_x(xb, T, si);
    st.consume_all();
    st.commit_rows(L.A[0], T, si);
    st.commit_x(L.x[0], T, si);
  }
  lds_barrier();
#pragma unroll 1
  for (int64_t p = 0; p < ntiles + 2; ++p) {
    const bool more = p + 1 < ntiles;
    const int nb = (int)((p + 1) % FNBUF);
    // staged registers live within one period (not carried across the loop)
    // the loads are unconditional (in the last periods they read the last tile
    // again, never committed): a conditional load is waited for at the merge
    Stage<ALIGNED, FS> st;
    st.load_rows(Ab, T, min(p + 1, ntiles - 1) * FS, si);
    st.load_x(xb, T, si);
    double ph[NP], z = 0.0;
    const int kk = lane & 15, c = wave * 4 + (lane >> 4);
    if (wave < FWD_PHI_WAVES && p < ntiles) {
      const int bb = (int)(p % FNBUF);
      phi_sweep<true>(L.A[bb], reinterpret_cast<const float*>(L.x[bb]), c, kk, ph, z);
      phi_store<true>(L.summ[p & 1][c], kk, ph, z);
    } else if (wave == FWD_CHAIN_WAVE && p >= 1 && p <= ntiles) {
      // on lanes 0-15 only: the other rows would repeat the chain and quadruple
      // its LDS reads of the link matrices
      if (lane < NP) chain<true, FC>(s, L.summ[(p - 1) & 1], L.carry[(p - 1) & 1], lane);
    } else if (wave == FWD_RERUN_WAVE && p >= 2) {
      const int64_t k = p - 2;
      const int bb = (int)(k % FNBUF);
      fwd_rerun<ALIGNED, FS>(L.A[bb], L.x[bb], L.carry[k & 1], yb, T, k * FS, lane);
    }
    st.consume_all();
    if (more) {
      st.commit_rows(L.A[nb], T, si);
      st.commit_x(L.x[nb], T, si);
    }
    lds_barrier();
  }
}

// LamT: the type of the adjoint rerun's recurrence (double: every output within
// a few 1e-8 of FP64 on the B=256, T=48000, N=16 inputs).
using LamT = double;

template <bool ALIGNED>
__global__ __launch_bounds__(THREADS) void tile_bwd(ChunkArgs<float> a) {
  __shared__ BwdLds L;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t b = blockIdx.x;
  const int64_t T = a.T;
  const int64_t ntiles = (T + S - 1) / S;
  const float* Ab = a.A + b * T * NP;
  const float* gb = a.x + b * T;
  // y is read for dA only; without it the window streams g again (never used)
  const float* yb = a.y ? a.y + b * T : gb;
  const float* zib = a.zi ? a.zi + b * NP : nullptr;
  float* dxb = a.dx ? a.dx + b * T : nullptr;
  float* dAb = a.dA ? a.dA + b * T * NP : nullptr;
  double s = 0.0;  // adjoint chain state (lanes 0-15 of the chain wave): zero above the signal
  const int si = tid;
  // tiles are taken from the last one down: period p starts tile ntiles-1-p
  {
    Stage<ALIGNED> st;
    const int64_t t0 = (ntiles - 1) * S;
    st.load_rows(Ab, T, t0, si);
    st.load_gy(gb, yb, zib, T, si);
    st.consume_all();
    st.commit_rows(L.A[0], T, si);
    st.commit_gy(L.g[0], L.y[0], zib, T, si);
  }
  lds_barrier();
#pragma unroll 1
  for (int64_t p = 0; p < ntiles + 3; ++p) {
    const bool more = p + 1 < ntiles;
    // staged registers live within one period; the loads are unconditional (in
    // the last periods they read tile 0 again, never committed)
    Stage<ALIGNED> st;
    st.load_rows(Ab, T, max(ntiles - 2 - p, (int64_t)0) * S, si);
    st.load_gy(gb, yb, zib, T, si);
    double ph[NP], z = 0.0;
    const int kk = lane & 15, c = bwd_phi_index(wave) * 4 + (lane >> 4);
    const int q = lane >> 2, ql = lane & 3;
    double P[4] = {0.0, 0.0, 0.0, 0.0};
    double now = 0.0;
    if (bwd_phi(wave)) {
      if (p < ntiles) phi_sweep<false>(L.A[p % NBUF], nullptr, c, kk, ph, z);
    } else if (wave == CHAIN_WAVE) {
      if (p >= 1 && p <= ntiles && lane < NP) chain<false>(s, L.summ, L.carry[(p - 1) & 1], lane);
    } else if (wave == DA_WAVE) {
      if (p >= 3) {
        const int64_t k = p - 3;
        grads<ALIGNED>(L.lam[k & 1], reinterpret_cast<const float*>(L.y[k % (NBUF + 1)]), dxb, dAb, T,
                       (ntiles - 1 - k) * S, lane);
      }
    } else {
      if (wave == ZS_WAVE && p < ntiles) {
        const int bb = (int)(p % NBUF);
        zero_state(BwdTile{L.A[bb], L.g[bb], L.y[p % (NBUF + 1)]}, P, now, q, ql);
      } else if (wave == RERUN_WAVE && p >= 2 && p <= ntiles + 1) {
        const int64_t k = p - 2;  // stage number of the tile
        const double (*cr)[NP] = L.carry[k & 1];
        LamT Pf[4];
        LamT nowf = (LamT)cr[q][0];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int rr = 4 * ql + 1 + m;
          Pf[(2 - m) & 3] = rr < NP ? (LamT)cr[q][rr] : (LamT)0;
        }
        const int bb = (int)(k % NBUF);
        lam_rerun(BwdTile{L.A[bb], L.g[bb], L.y[k % (NBUF + 1)]}, reinterpret_cast<float4*>(L.lam[k & 1]), Pf,
                  nowf, q, ql);
        if (k == ntiles - 1 && q == 0 && a.dzi) {
          // pending sums of positions -1-r: r = 0 in `now` (lane 0), 4ql+1+m in slot (2-m)&3
          float* dz = a.dzi + b * NP;
          if (ql == 0) dz[0] = (float)nowf;
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            const int rr = 4 * ql + 1 + m;
            if (rr < NP) dz[rr] = (float)Pf[(2 - m) & 3];
          }
        }
      }
    }
    lds_barrier();
    if (bwd_phi(wave) && p < ntiles) phi_store<false>(L.summ[c], kk, ph, z);
    if (wave == ZS_WAVE && p < ntiles) {
      double* w = &L.summ[q][NP * ROWP];
      if (ql == 0) w[0] = now;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int rr = 4 * ql + 1 + m;
        if (rr < NP) w[rr] = P[(2 - m) & 3];
      }
    }
    st.consume_all();
    if (more) {
      st.commit_rows(L.A[(p + 1) % NBUF], T, si);
      st.commit_gy(L.g[(p + 1) % NBUF], L.y[(p + 1) % (NBUF + 1)], zib, T, si);
    }
    lds_barrier();
  }
}

}  // namespace t16

// The per-signal strategy reads A, the per-sample streams and zi as float4 when
// every one of them (and dA) is 16-byte aligned; otherwise element by element.
static bool tile16_aligned(const ChunkArgs<float>& a) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  return al(a.A) && al(a.x) && (!a.y || al(a.y)) && (!a.dx || al(a.dx)) && (!a.dA || al(a.dA)) &&
         a.T % 4 == 0;
}

int launch_tile16_fwd(ChunkArgs<float> a, hipStream_t st) {
  if (a.B <= 0 || a.T <= 0) return 0;
  const dim3 grid((unsigned)a.B), block(t16::THREADS);
  if (tile16_aligned(a))
    hipLaunchKernelGGL(t16::tile_fwd<true>, grid, block, 0, st, a);
  else
    hipLaunchKernelGGL(t16::tile_fwd<false>, grid, block, 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_tile16_bwd(ChunkArgs<float> a, hipStream_t st) {
  if (a.B <= 0 || a.T <= 0) return 0;
  const dim3 grid((unsigned)a.B), block(t16::THREADS);
  if (tile16_aligned(a))
    hipLaunchKernelGGL(t16::tile_bwd<true>, grid, block, 0, st, a);
  else
    hipLaunchKernelGGL(t16::tile_bwd<false>, grid, block, 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // namespace tlpc
