// Chunked scan for order 16, float32 data (the B=256, T=48000, N=16 shape and
// any other float32 order-16 batch), built from the DPP blocks of tlpc_dpp16.h:
//
//   summary  forward: 16 chunks per workgroup, one 16-lane DPP row per chunk:
//            lane r sweeps the chunk backwards with the transposed recurrence
//            from a unit at output r, which gives row r of Phi (the pending
//            sums left before the chunk) and z[r] (the lambdas met on the way,
//            dotted with x); chunks 0..P-2.
//            adjoint: quads only (zero-state response of the transposed
//            recurrence -> w), chunks 1..P-1. The adjoint transition matrix of a
//            chunk is Phi^T: the transposed recurrence is the adjoint of the
//            forward one, so its homogeneous chunk map is the transpose. An
//            adjoint that is handed the forward's summaries skips Phi entirely;
//            otherwise it computes Phi with the summary kernel's rows.
//   carry    one wave per signal, float64.
//   rerun    one quad per chunk from its float64 carry, float64 arithmetic on the
//            float32 data (the rerun produces the outputs, so it keeps their
//            error at the final float32 rounding);
//            the adjoint rerun emits dx, the 64-byte dA row of every sample from
//            the quad's four lanes (one float4 each, a coalesced row store) and
//            dzi for chunk 0.
//
// Layouts: summ[b][c][r*16 + k] = Phi_c[r][k], summ[b][c][256 + r] = z_c[r]
// (c = 0..P-2); wsum[b][c-1][r] = w_c[r] (c = 1..P-1); carry[b][c][r].
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"

namespace tlpc {
namespace c16 {

using namespace dpp16;

constexpr int RERUN_THREADS = 256;

// ---------------------------------------------------------------------------
// Loads. Every load in the step loops is unconditional, at an index clamped into
// the signal; values that must read as zero are selected afterwards. A load under
// a branch leaves the compiler a phi to merge, and it waits for the load at the
// merge, which serialises the prefetch. Indices inside one signal are 32-bit: the
// dispatch keeps T * 16 below 2^31.
__device__ __forceinline__ float4 ld_row(const float* Ab, int t, int ql) {
  return reinterpret_cast<const float4*>(Ab + t * NP)[ql];
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Keeps the memory instructions of a step where the source puts them: ALU work
// may still move across, loads and stores may not. Without it the scheduler may
// gather a block's ring reloads at the end of the loop body, right before the
// next block consumes them, and the prefetch distance drops to nothing.
__device__ __forceinline__ void pin_loads() {
#ifndef TLPC_NO_PIN
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// The prologue fills of the prefetch rings pass through an empty asm statement.
// A ring slot is a loop phi whose incoming values are both loads (the prologue's
// and the previous iteration's); without the barrier InstCombine turns that phi
// of loads into one load at the loop header, issued right before its use, and
// the prefetch is gone.
__device__ __forceinline__ float opaque(float v) {
  asm("" : "+v"(v));
  return v;
}
__device__ __forceinline__ float4 opaque(float4 v) {
  asm("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
  return v;
}
// A loop-carried state that comes from memory (the carries) passes through here
// before the loop, which makes the kernel wait for it once there. Otherwise the
// wait lands at its first use inside the loop, where the wait-count pass merges
// it with the back edge and repeats it every block: a wait that also drains the
// ring loads of the previous block.
template <int K>
__device__ __forceinline__ void settle(double (&v)[K]) {
#pragma unroll
  for (int i = 0; i < K; ++i) asm volatile("" : "+v"(v[i]));
}

// y_u of one signal, with the initial conditions y_{-1-k} = zi[k] (k < 16) and
// zero elsewhere before 0: the address is always valid (indices past T clamp to
// T-1, a value the callers never use) and yok says whether the value counts. The
// two are applied at different steps so that the load is never made conditional.
__device__ __forceinline__ const float* yaddr(const float* yb, const float* zib, int u, int T) {
  return u >= 0 ? yb + min(u, T - 1) : (zib ? zib + min(-u - 1, NP - 1) : yb);
}
__device__ __forceinline__ bool yok(const float* zib, int u) {
  return u >= 0 || (zib != nullptr && -u - 1 < NP);
}

// Workgroup barrier that orders LDS only. __syncthreads() also fences global
// memory, which waits for every load in flight (the prefetch rings); the waves
// here exchange data through LDS alone.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int K>
__device__ __forceinline__ void put(float4& o, int ql, float v) {
  // component K & 3 of lane K >> 2 of the quad
  const bool mine = ql == (K >> 2);
  if constexpr ((K & 3) == 0) o.x = mine ? v : o.x;
  if constexpr ((K & 3) == 1) o.y = mine ? v : o.y;
  if constexpr ((K & 3) == 2) o.z = mine ? v : o.z;
  if constexpr ((K & 3) == 3) o.w = mine ? v : o.w;
}

// Stores the 16 samples [tb, tb+16) that a quad gathered (lane ql holds
// tb+4ql .. tb+4ql+3): one float4 per lane when the block is whole and aligned.
__device__ __forceinline__ void store16(float* vb, int tb, int T, int ql, float4 o) {
  float* p = vb + tb + 4 * ql;
  if (tb + 16 <= T && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    *reinterpret_cast<float4*>(p) = o;
  } else {
    const int u = tb + 4 * ql;
    if (u < T) p[0] = o.x;
    if (u + 1 < T) p[1] = o.y;
    if (u + 2 < T) p[2] = o.z;
    if (u + 3 < T) p[3] = o.w;
  }
}

// ---------------------------------------------------------------------------
// Per-sample scalar streams (x, g, y). With VEC (T a multiple of 4 and 16-byte
// aligned rows) lane ql of a quad loads samples [tb+4ql, tb+4ql+4) of a 16-step
// block as one float4, a block ahead, and each step takes its sample from the
// owning lane by a quad broadcast: one 64-byte request per quad per 16 steps
// instead of one per step. Without VEC every step loads its sample.
constexpr int qbcast_ctrl(int q) { return q * 0x55; }  // quad_perm [q,q,q,q]

template <int K>
__device__ __forceinline__ float comp(const float4& v) {
  if constexpr ((K & 3) == 0) return v.x;
  if constexpr ((K & 3) == 1) return v.y;
  if constexpr ((K & 3) == 2) return v.z;
  return v.w;
}
// sample K (0..15) of a block held as one float4 per quad lane, in every lane
template <int K>
__device__ __forceinline__ float qsel(const float4& v) {
  return qperm<qbcast_ctrl(K >> 2)>(comp<K>(v));
}

// float4 of samples [base+4ql, base+4ql+4) of a row of length T (T % 4 == 0);
// out-of-range blocks read a valid float4 whose values the callers never use.
__device__ __forceinline__ float4 ld4(const float* p, int base, int ql, int T) {
  const int i = clampi(base + 4 * ql, 0, T - 4);
  return *reinterpret_cast<const float4*>(p + i);
}

// ---------------------------------------------------------------------------
// Forward quad run over [t0, t0+len): 16-step blocks, a PF-deep ring of rows and
// inputs. Steps past the end of the signal run on clamped data and store nothing.
struct FwdIo {
  const float* Ab;  // A of the signal, [T][16]
  const float* xb;
  float* yb;
  int T;
};

// Whole-block ring reloads go out RG rows at a time (RG * 64 contiguous bytes per
// quad): a stream's consecutive cache lines then reach the memory controller
// together instead of a step apart.
#ifndef TLPC_RG
#define TLPC_RG 4
#endif
constexpr int RG = TLPC_RG;
static_assert(PF % RG == 0, "reload groups tile the ring");

// WHOLE: every 16-step block that the run uses lies wholly inside the signal
// (the run covers whole blocks: a full chunk, or T a multiple of 16). The ring
// reloads are then immediate offsets from one base per block (nb: lane ql's
// float4 of the next block's first row), clamped once per block so that the
// prefetch past the run stays in the signal; otherwise every step clamps its row.
template <int JJ, bool STORE, bool VEC, bool WHOLE, typename V>
__device__ __forceinline__ void fwd_q16(V (&R)[4], float4 (&cq)[PF], float (&xq)[PF], float4& out,
                                        const float4& xc, const FwdIo& io, int tb, int ql,
                                        const float4* nb) {
  if constexpr (JJ < 16) {
    // the step consumes its ring slot before the slot is reloaded (for step
    // JJ + PF), so the old and new values never overlap and the slot stays in
    // one register across iterations (no copies, no wait at the loop latch)
    float xv;
    if constexpr (VEC) xv = qsel<JJ>(xc);
    else xv = xq[JJ % PF];
    const V y = fwd_quad_step<JJ & 3>(R, cq[JJ % PF], xv, ql);
    const int tn = min(tb + JJ + PF, io.T - 1);
    if constexpr (WHOLE) {
      // RG consumed slots at a time: RG consecutive rows, issued together
      if constexpr (JJ % RG == RG - 1) {
#pragma unroll
        for (int k = JJ - RG + 1; k <= JJ; ++k) cq[k] = nb[k * (NP / 4)];
      }
    } else {
      cq[JJ % PF] = ld_row(io.Ab, tn, ql);
    }
    if constexpr (!VEC) xq[JJ % PF] = io.xb[tn];
    pin_loads();
    if constexpr (STORE) put<JJ>(out, ql, (float)y);
    fwd_q16<JJ + 1, STORE, VEC, WHOLE>(R, cq, xq, out, xc, io, tb, ql, nb);
  }
}

template <bool STORE, bool VEC, bool WHOLE, typename V>
__device__ __forceinline__ void fwd_quad_run(V (&R)[4], const FwdIo& io, int t0, int len, int ql) {
  float4 cq[PF];
  float xq[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int tj = min(t0 + j, io.T - 1);
    cq[j] = opaque(ld_row(io.Ab, tj, ql));
    if constexpr (!VEC) xq[j] = opaque(io.xb[tj]);
  }
  float4 xc = make_float4(0.f, 0.f, 0.f, 0.f);
  if constexpr (VEC) xc = opaque(ld4(io.xb, t0, ql, io.T));
  settle(R);
  for (int jb = 0; jb < len; jb += 16) {
    const int tb = t0 + jb;
    float4 xn = xc;
    if constexpr (VEC) xn = ld4(io.xb, tb + 16, ql, io.T);  // used by the next block
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4* nb = nullptr;
    if constexpr (WHOLE) nb = reinterpret_cast<const float4*>(io.Ab + min(tb + 16, io.T - 16) * NP) + ql;
    fwd_q16<0, STORE, VEC, WHOLE>(R, cq, xq, out, xc, io, tb, ql, nb);
    if constexpr (STORE) store16(io.yb, tb, io.T, ql, out);
    xc = xn;
  }
}

// ---------------------------------------------------------------------------
// Adjoint quad run, descending over the 16-step blocks of [t0, t0+L). Steps at
// or past T are no-ops of a zero state (zero coefficients and gradient). With
// OUT it also emits dx (gathered per block), the dA rows (when DA) and keeps the
// window Y of y_{t-1-4ql-m} for them.
struct AdjIo {
  const float* Ab;
  const float* gb;
  const float* yb;
  const float* zib;
  float* dxb;
  float* dAb;
  int T;
};

// VEC blocks of the adjoint: g of the current block, and the y samples that
// enter the window (y_{t-17} at step t): block tb-16 for the first 15 steps of
// block tb, block tb-32 for its last step.
struct AdjVec {
  float4 g, ycur, ynext;
};

// WHOLE as for the forward run, descending: pb is lane ql's float4 of the first
// row of block tb - 16 (clamped into the signal), and with T a multiple of 16 a
// block is either inside the signal or wholly past its end (no live step).
template <int JJ, bool OUT, bool DA, bool VEC, bool WHOLE>
__device__ __forceinline__ void adj_q16(double (&P)[4], double& now, float (&Y)[4],
                                        float4 (&cq)[PF], float (&gq)[PF], float (&yq)[PF],
                                        float4& dxo, const AdjVec& vb, const AdjIo& io, int tb,
                                        int ql, const float4* pb) {
  if constexpr (JJ < 16) {
    constexpr int J = (3 - JJ) & 3;  // slot of t = tb + 15 - JJ (tb is a multiple of 16)
    const int t = tb + 15 - JJ;
    float gv, yraw = 0.f;
    if constexpr (VEC) {
      gv = qsel<15 - JJ>(vb.g);
      if constexpr (DA) yraw = (JJ < 15) ? qsel<(14 - JJ) & 15>(vb.ycur) : qsel<15>(vb.ynext);
    } else {
      gv = gq[JJ % PF];
      yraw = yq[JJ % PF];
    }
    const bool live = t < io.T;
    // ring slots are consumed before they are reloaded (see fwd_q16)
    const double lam = adj_quad_step<J>(P, now, cq[JJ % PF], gv, ql, live);
    {
      const int tc = clampi(t - PF, 0, io.T - 1);
      if constexpr (WHOLE) {
        if constexpr (JJ % RG == RG - 1) {
#pragma unroll
          for (int k = JJ - RG + 1; k <= JJ; ++k) cq[k] = pb[(15 - k) * (NP / 4)];
        }
      } else {
        cq[JJ % PF] = ld_row(io.Ab, tc, ql);
      }
      if constexpr (!VEC) {
        gq[JJ % PF] = io.gb[tc];
        if constexpr (DA) yq[JJ % PF] = *yaddr(io.yb, io.zib, t - PF - 17, io.T);
      }
    }
    pin_loads();
    if constexpr (OUT) {
      const float lf = (float)lam;
      put<15 - JJ>(dxo, ql, lf);
      if constexpr (DA) {
        if (live) {
          float4 d;
          d.x = -lf * Y[(J - 1) & 3];
          d.y = -lf * Y[(J - 2) & 3];
          d.z = -lf * Y[(J - 3) & 3];
          d.w = -lf * Y[(J - 4) & 3];
          reinterpret_cast<float4*>(io.dAb + t * NP)[ql] = d;
        }
        constexpr int S = (J - 1) & 3;
        const float down = qperm<QP_DOWN>(Y[S]);
        // VEC blocks already hold zi (or zero) before time 0
        const float ynew = (VEC || yok(io.zib, t - 17)) ? yraw : 0.f;
        Y[S] = (ql == 3) ? ynew : down;  // lane 3 takes y_{t-17}
      }
    }
    adj_q16<JJ + 1, OUT, DA, VEC, WHOLE>(P, now, Y, cq, gq, yq, dxo, vb, io, tb, ql, pb);
  }
}

// y block [v, v+16) as one float4 per lane, for a block start v (a multiple of
// 16) at or after -32: v >= 0 loads (clamped), v == -16 is the initial state
// reversed (y_{-1-k} = zi[k], zrev precomputed), v == -32 is zero.
__device__ __forceinline__ float4 yblock_fix(const float4& raw, const float4& zrev, int v) {
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4& alt = (v == -16) ? zrev : z;
  float4 r;
  r.x = v >= 0 ? raw.x : alt.x;
  r.y = v >= 0 ? raw.y : alt.y;
  r.z = v >= 0 ? raw.z : alt.z;
  r.w = v >= 0 ? raw.w : alt.w;
  return r;
}

// One 16-step block of the adjoint run. The float4 blocks of g and y live in
// two-slot rings indexed by the block's parity PAR (block k of the run, k = 0 at
// the top): gs[PAR] holds g of this block, ys[PAR ^ 1] the y block below it and
// ys[PAR] the one below that (read by the last step only). Each slot is reloaded
// in place as soon as its last reader is done, for the block two (g) or three (y)
// further down; with the parity a template argument no value moves between
// registers at the loop latch, and the latch needs no wait for loads in flight.
template <int PAR, bool OUT, bool DA, bool VEC, bool WHOLE>
__device__ __forceinline__ void adj_block(double (&P)[4], double& now, float (&Y)[4],
                                          float4 (&cq)[PF], float (&gq)[PF], float (&yq)[PF],
                                          float4 (&gs)[2], float4 (&ys)[2], const float4& zrev,
                                          const AdjIo& io, int tb, int ql) {
  AdjVec vb;
  if constexpr (VEC) {
    vb.g = gs[PAR];
    if constexpr (DA) {
      vb.ycur = yblock_fix(ys[PAR ^ 1], zrev, tb - 16);
      vb.ynext = yblock_fix(ys[PAR], zrev, tb - 32);
    }
  }
  float4 dxo = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4* pb = nullptr;
  if constexpr (WHOLE) pb = reinterpret_cast<const float4*>(io.Ab + clampi(tb - 16, 0, io.T - 16) * NP) + ql;
  adj_q16<0, OUT, DA, VEC, WHOLE>(P, now, Y, cq, gq, yq, dxo, vb, io, tb, ql, pb);
  if constexpr (OUT) {
    if (io.dxb) store16(io.dxb, tb, io.T, ql, dxo);
  }
  if constexpr (VEC) {
    gs[PAR] = ld4(io.gb, tb - 32, ql, io.T);
    if constexpr (DA) ys[PAR ^ 1] = ld4(io.yb, tb - 48, ql, io.T);
  }
}

// The run takes its blocks in pairs (parities 0, 1); an odd block count ends
// with one more parity-0 block.
template <bool OUT, bool DA, bool VEC, bool WHOLE>
__device__ __forceinline__ void adj_quad_run(double (&P)[4], double& now, const AdjIo& io, int t0,
                                             int L, int ql) {
  const int tstart = t0 + L - 1;
  // y window: slot u mod 4 holds y_u, u in [t-4-4ql, t-1-4ql] for t = tstart
  // (values past T never meet a nonzero lambda)
  float Y[4] = {0.f, 0.f, 0.f, 0.f};
  if constexpr (DA) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int u = tstart - 1 - 4 * ql - m;
      const float v = *yaddr(io.yb, io.zib, u, io.T);
      Y[(2 - m) & 3] = yok(io.zib, u) ? v : 0.f;
    }
  }
  float4 cq[PF];
  float gq[PF], yq[PF];
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int tc = clampi(tstart - j, 0, io.T - 1);
    cq[j] = opaque(ld_row(io.Ab, tc, ql));
    if constexpr (!VEC) {
      gq[j] = opaque(io.gb[tc]);
      yq[j] = DA ? opaque(*yaddr(io.yb, io.zib, tstart - j - 17, io.T)) : 0.f;
    }
  }
  const int tb0 = t0 + L - 16;
  float4 gs[2], ys[2], zrev = make_float4(0.f, 0.f, 0.f, 0.f);
  gs[0] = gs[1] = ys[0] = ys[1] = zrev;
  if constexpr (VEC) {
    gs[0] = opaque(ld4(io.gb, tb0, ql, io.T));
    gs[1] = opaque(ld4(io.gb, tb0 - 16, ql, io.T));
    if constexpr (DA) {
      ys[1] = opaque(ld4(io.yb, tb0 - 16, ql, io.T));
      ys[0] = opaque(ld4(io.yb, tb0 - 32, ql, io.T));
      // zi as y_{-16..-1}: lane ql, component m holds y_{-16+4ql+m} = zi[15-4ql-m]
      if (io.zib) {
        const float4 zv = *reinterpret_cast<const float4*>(io.zib + 4 * ql);
        constexpr int REV = 0x1B;  // quad_perm [3,2,1,0]
        zrev.x = qperm<REV>(zv.w);
        zrev.y = qperm<REV>(zv.z);
        zrev.z = qperm<REV>(zv.y);
        zrev.w = qperm<REV>(zv.x);
      }
    }
  }
  int tb = tb0;
  for (; tb - 16 >= t0; tb -= 32) {
    adj_block<0, OUT, DA, VEC, WHOLE>(P, now, Y, cq, gq, yq, gs, ys, zrev, io, tb, ql);
    adj_block<1, OUT, DA, VEC, WHOLE>(P, now, Y, cq, gq, yq, gs, ys, zrev, io, tb - 16, ql);
  }
  if (tb >= t0) adj_block<0, OUT, DA, VEC, WHOLE>(P, now, Y, cq, gq, yq, gs, ys, zrev, io, tb, ql);
}

// ---------------------------------------------------------------------------
// Forward summaries: chunks c = 0..P-2 (full length L), 16 per workgroup, one
// 16-lane DPP row per chunk.
//
// Lane r of a row computes row r of Phi and z[r] together, by sweeping the
// chunk backwards with the transposed recurrence from a unit at output r (the
// sample t0+L-1-r): the pending sums left for the positions before the chunk are
// d y_{t0+L-1-r} / d y_{t0-1-k} = Phi[r][k], and the lambdas met on the way are
// d y_{t0+L-1-r} / d x_t, so z[r] = sum_t lambda_t x_t. One pass over A serves
// both (the forward form needs a 17th, zero-state, recurrence beside the 16
// basis responses), each step's 16 products are independent of each other, and
// the sweep ends on the chunk's first rows, which the forward rerun reads first.
//
// Lane kk holds A[t, kk+1] of the step (a 16-slot ring, consumed then reloaded
// for the block below: the row base np is clamped once per block so the last
// prefetch stays in the signal) and x of step tb+kk of the block.
// Steps 4M+3 .. 4M of a 16-step block from ring slot M: the float4 of rows
// tb+4M .. tb+4M+3 (lane kk holds row tb+4M+(kk>>2), coefficients 4(kk&3) ..
// 4(kk&3)+3). The slot is converted, then reloaded with the block below.
template <int J, bool WITH_Z>
__device__ __forceinline__ void adj_row_quad(double (&p)[NP], double& z, const double (&d)[4], double xv) {
  if constexpr (J >= 0) {
    const double lam = adj_row_step4<J>(p, d);
    if constexpr (WITH_Z) fma_bcast<J>(z, xv, lam);
    if constexpr ((J & 3) != 0) adj_row_quad<J - 1, WITH_Z>(p, z, d, xv);
  }
}
template <int M, bool WITH_Z>
__device__ __forceinline__ void adj_row_block(double (&p)[NP], double& z, float4 (&cq)[4], double xv,
                                              const float4* nq) {
  if constexpr (M >= 0) {
    const double d[4] = {(double)cq[M].x, (double)cq[M].y, (double)cq[M].z, (double)cq[M].w};
    cq[M] = nq[M * 16];
    pin_loads();
    adj_row_quad<4 * M + 3, WITH_Z>(p, z, d, xv);
    pin_loads();
    adj_row_block<M - 1, WITH_Z>(p, z, cq, xv, nq);
  }
}

template <bool WITH_Z>
__global__ __launch_bounds__(256) void sum_fwd(ChunkArgs<float> a) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t cps = a.P - 1;
  const int64_t nch = a.B * cps;
  const int64_t cid = (int64_t)blockIdx.x * 16 + wave * 4 + (lane >> 4);
  if (cid >= nch) return;
  const int L = a.L;
  const int kk = lane & 15;
  const int64_t b = cid / cps;
  const int c = (int)(cid - b * cps);
  const int t0 = c * L;
  // lane kk reads float4 kk of a 4-row group: row (kk >> 2), coefficients 4(kk & 3)..
  const float4* ap = reinterpret_cast<const float4*>(a.A + (b * a.T + t0) * NP) + kk;
  const float* xp = a.x + b * a.T + t0 + kk;
  double p[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) p[i] = (i == NP - 1 - kk) ? 1.0 : 0.0;  // unit at t0+L-1-kk
  double z = 0.0;
  float4 cq[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) cq[m] = opaque(ap[(L - NP + 4 * m) * 4]);
  float xf = WITH_Z ? opaque(xp[L - NP]) : 0.f;
  for (int jb = L - NP; jb >= 0; jb -= NP) {
    const int jn = jb >= NP ? jb - NP : 0;  // the block below (block 0 again after the last)
    const double xv = (double)xf;
    float xn = 0.f;
    if constexpr (WITH_Z) xn = xp[jn];
    adj_row_block<3, WITH_Z>(p, z, cq, xv, ap + jn * 4);
    xf = xn;
  }
  double* out = a.summ + cid * (NP * NP + NP);
#pragma unroll
  for (int k = 0; k < NP; ++k) out[kk * NP + k] = p[NP - 1 - k];  // Phi[kk][k]
  if constexpr (WITH_Z) out[NP * NP + kk] = z;
}

// Adjoint zero-state responses: chunks c = 1..P-1, one quad each (the last may
// be partial; its steps past the end are no-ops of the zero state).
constexpr int WSUM_THREADS = 256;
// Stream modes of the kernels below: 0 per-step loads, 1 float4 blocks of the
// per-sample streams (VEC), 2 VEC plus whole-block A rows (T a multiple of 16).
template <int MODE>
__global__ __launch_bounds__(WSUM_THREADS) void sum_adj(ChunkArgs<float> a) {
  const int64_t cps = a.P - 1;
  const int64_t cid = ((int64_t)blockIdx.x * WSUM_THREADS + threadIdx.x) >> 2;
  if (cid >= a.B * cps) return;
  const int ql = threadIdx.x & 3;
  const int64_t b = cid / cps;
  const int c = (int)(cid - b * cps) + 1;
  double P[4] = {0.0, 0.0, 0.0, 0.0};
  double now = 0.0;
  const AdjIo io{a.A + b * a.T * NP, a.x + b * a.T, nullptr, nullptr, nullptr, nullptr, (int)a.T};
  adj_quad_run<false, false, (MODE >= 1), (MODE == 2)>(P, now, io, c * a.L, a.L, ql);
  // w[0] = now (lane 0); w[4ql+1+m] = P[(2-m) & 3]
  double* out = a.wsum + cid * NP;
  if (ql == 0) out[0] = now;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int rr = 4 * ql + 1 + m;
    if (rr < NP) out[rr] = P[(2 - m) & 3];
  }
}

// ---------------------------------------------------------------------------
// Forward rerun: one quad per chunk (all P chunks of all signals).
// The reruns walk the chunks from the last one down: the A rows that the
// summary kernel before them read last are then read first, while they may
// still be in the last-level (MALL) cache.
__device__ __forceinline__ int64_t rerun_chunk(const ChunkArgs<float>& a) {
  return a.B * a.P - 1 - (((int64_t)blockIdx.x * RERUN_THREADS + threadIdx.x) >> 2);
}

template <int MODE>
__global__ __launch_bounds__(RERUN_THREADS) void rerun_fwd(ChunkArgs<float> a) {
  const int64_t cid = rerun_chunk(a);
  if (cid < 0) return;
  const int ql = threadIdx.x & 3;
  const int64_t b = cid / a.P;
  const int c = (int)(cid - b * a.P);
  const int T = (int)a.T;
  const int t0 = c * a.L;
  const int len = min(T - t0, a.L);
  const double* cr = a.carry + cid * NP;
  double R[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) R[3 - m] = cr[4 * ql + m];  // y_{t0-1-4ql-m}, slot 3-m
  const FwdIo io{a.A + b * a.T * NP, a.x + b * a.T, a.y + b * a.T, T};
  fwd_quad_run<true, (MODE >= 1), (MODE == 2)>(R, io, t0, len, ql);
}

// Adjoint rerun: one quad per chunk, descending from the chunk's last step (steps
// past the end of the last chunk are no-ops of its zero carry).
template <bool DA, int MODE>
__global__ __launch_bounds__(RERUN_THREADS) void rerun_adj(ChunkArgs<float> a) {
  const int64_t cid = rerun_chunk(a);
  if (cid < 0) return;
  const int ql = threadIdx.x & 3;
  const int64_t b = cid / a.P;
  const int c = (int)(cid - b * a.P);
  const int T = (int)a.T;
  const int L = a.L;
  // pending state from the carry: now = p_0, window slot (2-m)&3 = p_{4ql+1+m}
  const double* cr = a.carry + cid * NP;
  double P[4];
  double now = cr[0];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int rr = 4 * ql + 1 + m;
    P[(2 - m) & 3] = rr < NP ? cr[rr] : 0.0;
  }
  const AdjIo io{a.A + b * a.T * NP,
                 a.x + b * a.T,
                 a.y + b * a.T,
                 a.zi ? a.zi + b * NP : nullptr,
                 a.dx ? a.dx + b * a.T : nullptr,
                 a.dA ? a.dA + b * a.T * NP : nullptr,
                 T};
  adj_quad_run<true, DA, (MODE >= 1), (MODE == 2)>(P, now, io, c * L, L, ql);
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
// Carry chain, one workgroup per signal: s_{l+1} = M_l s_l + v_l over the P-1
// links. Forward: link l is chunk l, M = Phi_l, v = z_l. Adjoint: link l is
// chunk P-1-l, M = Phi_{P-1-l}^T, v = w_{P-1-l}; the first link (the last chunk)
// meets the zero state past the end of the signal, so it needs no matrix.
//
// Wave 0 chains; the other LOADERS waves load. Links come in batches of CB;
// batch j is loaded by wave 1 + j % LOADERS, which issues its global loads
// DEPTH = LOADERS batches before the chain needs it and writes it to one of two
// LDS buffers in the iteration just before. One barrier per batch keeps the
// waves in step, so the chain never waits for memory once the pipeline is full.
constexpr int CB = 8;
constexpr int SUMM = NP * NP + NP;
// LDS layout of one link: 16 matrix rows padded to 17 doubles (row r at r*17,
// so that 16 lanes reading one column, or one row, hit 16 distinct bank pairs),
// then the 16-vector.
constexpr int ROWP = NP + 1;
constexpr int LINKP = NP * ROWP + NP;
constexpr int CARRY_THREADS = 256;
constexpr int LOADERS = CARRY_THREADS / 64 - 1;
constexpr int DEPTH = LOADERS;

template <bool FWD>
__global__ __launch_bounds__(CARRY_THREADS) void carry16(ChunkArgs<float> a) {
  __shared__ double buf[2][CB][LINKP];
  const int64_t b = blockIdx.x;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t P = a.P;
  const int64_t links = P - 1;
  const int64_t nbatch = (links + CB - 1) / CB;
  const double* summ = a.summ + b * links * SUMM;
  const double* wsum = FWD ? nullptr : a.wsum + b * links * NP;
  double* carry = a.carry + b * P * NP;
  constexpr int PL = (SUMM + 63) / 64;  // doubles per loader lane per link
  double stage[CB][PL];
  // A loader lane takes elements lane, lane+64, ... of each link of its batch:
  // one base per link plus a constant, no division by the link size. Loads are
  // unconditional, at a link index clamped into the chain and an element index
  // clamped into the link (the chain never reads those slots), with the
  // adjoint's choices made on the address; a load under a branch would be
  // waited for at once.
  auto fetch = [&](int64_t j) {
    const int64_t l0 = j * CB;
#pragma unroll
    for (int li = 0; li < CB; ++li) {
      const int64_t l = min(l0 + li, links - 1);
      if constexpr (FWD) {
        const double* src = summ + l * SUMM;
#pragma unroll
        for (int q = 0; q < PL; ++q) stage[li][q] = src[min(q * 64 + lane, SUMM - 1)];
      } else {
        // link l is chunk P-1-l: Phi^T of the forward summary of that chunk (link
        // 0, the last chunk, meets the zero state and has no matrix), then its w
        const double* ms = summ + (links - (l > 0 ? l : 1)) * SUMM;
        const double* ws = wsum + (links - 1 - l) * NP;
#pragma unroll
        for (int q = 0; q < PL; ++q) {
          const int off = q * 64 + lane;
          const bool w = off >= NP * NP;
          const double raw = w ? ws[min(off - NP * NP, NP - 1)] : ms[off];
          stage[li][q] = (w || l > 0) ? raw : 0.0;
        }
      }
    }
  };
  auto commit = [&](int64_t j) {
    double(*dst)[LINKP] = buf[j & 1];
#pragma unroll
    for (int li = 0; li < CB; ++li) {
#pragma unroll
      for (int q = 0; q < PL; ++q) {
        const int off = q * 64 + lane;
        if (off < SUMM) dst[li][off < NP * NP ? off + off / NP : off + NP] = stage[li][q];
      }
    }
  };
  double s = 0.0;
  if (wave == 0 && FWD && lane < NP && a.zi) s = (double)a.zi[b * NP + lane];
  const int row = lane & 15;
  for (int64_t i = -DEPTH; i < nbatch; ++i) {
    if (wave > 0) {
      const int64_t jn = i + DEPTH;  // issue
      if (jn < nbatch && 1 + jn % LOADERS == wave) fetch(jn);
      const int64_t jc = i + 1;  // commit (loaded DEPTH - 1 iterations ago)
      if (jc >= 0 && jc < nbatch && 1 + jc % LOADERS == wave) commit(jc);
    } else if (i >= 0) {
      const int64_t l0 = i * CB;
      const int nb = (int)((links - l0) < CB ? (links - l0) : CB);
      const double(*src)[LINKP] = buf[i & 1];
      // the LDS reads of link li + 1 go out before the chain step of link li, so
      // the chain waits for LDS only once per batch (links past nb read slots of
      // the batch buffer that nothing uses)
      double m[2][NP], zv[2];
      auto ldm = [&](int li, double(&mm)[NP], double& z) {
#pragma unroll
        for (int k = 0; k < NP; ++k) mm[k] = FWD ? src[li][row * ROWP + k] : src[li][k * ROWP + row];
        z = src[li][NP * ROWP + row];
      };
      ldm(0, m[0], zv[0]);
#pragma unroll
      for (int li = 0; li < CB; ++li) {
        if (li + 1 < CB) ldm(li + 1, m[(li + 1) & 1], zv[(li + 1) & 1]);
        if (li < nb) {
          const int64_t l = l0 + li;
          if (lane < NP) carry[(FWD ? l : P - 1 - l) * NP + lane] = s;
          s = chain_link(s, m[li & 1], zv[li & 1]);
        }
      }
    }
    lds_barrier();
  }
  if (wave == 0 && lane < NP) carry[(FWD ? P - 1 : 0) * NP + lane] = s;
}

}  // namespace c16

int launch_chunk16_carry(ChunkArgs<float> a, bool fwd, hipStream_t st) {
  if (fwd)
    hipLaunchKernelGGL(c16::carry16<true>, dim3((unsigned)a.B), dim3(c16::CARRY_THREADS), 0, st, a);
  else
    hipLaunchKernelGGL(c16::carry16<false>, dim3((unsigned)a.B), dim3(c16::CARRY_THREADS), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

// The per-sample streams are read as float4 blocks when every row starts on a
// 16-byte boundary: T a multiple of 4 and 16-byte aligned x / g, y and zi.
static bool vec_ok(const ChunkArgs<float>& a) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  return a.T % 4 == 0 && al(a.x) && (!a.y || al(a.y)) && (!a.zi || al(a.zi));
}
// Stream mode of the quad kernels: whole 16-step blocks need T % 16 == 0 (chunk
// starts are multiples of 16), so every block is inside the signal or past it.
static int stream_mode(const ChunkArgs<float>& a) {
  if (!vec_ok(a)) return 0;
  return a.T % 16 == 0 ? 2 : 1;
}

int launch_chunk16_fwd_summary(ChunkArgs<float> a, bool with_z, hipStream_t st) {
  const int64_t n = a.B * (a.P - 1);
  const dim3 grid((unsigned)((n + 15) / 16));
  if (n > 0 && with_z)
    hipLaunchKernelGGL(c16::sum_fwd<true>, grid, dim3(256), 0, st, a);
  else if (n > 0)
    hipLaunchKernelGGL(c16::sum_fwd<false>, grid, dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}
int launch_chunk16_adj_summary(ChunkArgs<float> a, hipStream_t st) {
  const int64_t lanes = a.B * (a.P - 1) * 4;
  const dim3 grid((unsigned)((lanes + c16::WSUM_THREADS - 1) / c16::WSUM_THREADS));
  if (lanes <= 0) return 0;
  switch (stream_mode(a)) {
    case 2: hipLaunchKernelGGL(c16::sum_adj<2>, grid, dim3(c16::WSUM_THREADS), 0, st, a); break;
    case 1: hipLaunchKernelGGL(c16::sum_adj<1>, grid, dim3(c16::WSUM_THREADS), 0, st, a); break;
    default: hipLaunchKernelGGL(c16::sum_adj<0>, grid, dim3(c16::WSUM_THREADS), 0, st, a); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}
int launch_chunk16_fwd_rerun(ChunkArgs<float> a, hipStream_t st) {
  const int64_t lanes = a.B * a.P * 4;
  const dim3 grid((unsigned)((lanes + c16::RERUN_THREADS - 1) / c16::RERUN_THREADS));
  switch (stream_mode(a)) {
    case 2: hipLaunchKernelGGL(c16::rerun_fwd<2>, grid, dim3(c16::RERUN_THREADS), 0, st, a); break;
    case 1: hipLaunchKernelGGL(c16::rerun_fwd<1>, grid, dim3(c16::RERUN_THREADS), 0, st, a); break;
    default: hipLaunchKernelGGL(c16::rerun_fwd<0>, grid, dim3(c16::RERUN_THREADS), 0, st, a); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}
int launch_chunk16_adj_rerun(ChunkArgs<float> a, hipStream_t st) {
  const int64_t lanes = a.B * a.P * 4;
  const dim3 grid((unsigned)((lanes + c16::RERUN_THREADS - 1) / c16::RERUN_THREADS));
  const int m = stream_mode(a);
  if (a.dA && m == 2)
    hipLaunchKernelGGL((c16::rerun_adj<true, 2>), grid, dim3(c16::RERUN_THREADS), 0, st, a);
  else if (a.dA && m == 1)
    hipLaunchKernelGGL((c16::rerun_adj<true, 1>), grid, dim3(c16::RERUN_THREADS), 0, st, a);
  else if (a.dA)
    hipLaunchKernelGGL((c16::rerun_adj<true, 0>), grid, dim3(c16::RERUN_THREADS), 0, st, a);
  else if (m == 2)
    hipLaunchKernelGGL((c16::rerun_adj<false, 2>), grid, dim3(c16::RERUN_THREADS), 0, st, a);
  else if (m == 1)
    hipLaunchKernelGGL((c16::rerun_adj<false, 1>), grid, dim3(c16::RERUN_THREADS), 0, st, a);
  else
    hipLaunchKernelGGL((c16::rerun_adj<false, 0>), grid, dim3(c16::RERUN_THREADS), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // namespace tlpc
