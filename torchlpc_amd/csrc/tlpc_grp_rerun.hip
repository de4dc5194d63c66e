// Reruns (phase 3 of the chunked scan) with prefetched streams and several
// lanes per chunk, for float32 time-varying filters of order 8, 16 and 32 in
// uniform or ragged batches (orders 2 and 4: tlpc_quad_rerun.hip). Forward:
// G = NP / 4 lanes share one chunk's serial recurrence, lane gl owning lags
// 4gl+1 .. 4gl+4, so the G lanes of a chunk load a coefficient row as G float4
// (one 16 * G-byte piece) and a wave's row loads cover 64 / G rows instead of
// 64. Adjoint: NP / 2 lanes with two lags each. This is the quad
// scheme of the order-16 kernels (tlpc_dpp16.h) with the group width a template
// parameter: partial sums meet through DPP (quad_perm within quads,
// row_half_mirror across the two quads of an 8-lane group), lambda leaves the
// group's lane 0 the same way, and the state windows move between neighbouring
// lanes through row_shr:1 / row_shl:1.
//
// Arithmetic is float64 on the float32 data, as in the one-lane reruns, so the
// only single-precision rounding of an output is its store. Rows, samples and
// (adjoint) the entering y values are prefetched RING steps ahead at clamped
// times; steps past the end of a signal compute on clamped data and store what
// the last live step stored (forward) or have lambda = 0 (adjoint).
//
// Reference recurrence: torchlpc/csrc/scan_cpu.cpp:112-120; adjoint:
// torchlpc/core.py:186-229.
#include "tlpc_dpp16.h"
#include "tlpc_kernels.h"
#include "tlpc_vmem.h"

namespace tlpc {
namespace grp {

using dpp16::qperm;
using namespace vmem;

// Forward group shape: G lanes per chunk, RING steps in flight per lane (a
// multiple of 4).
template <int NP>
struct Grp {
  static constexpr int G = NP / 4;
  static constexpr int RING = 8;
};
constexpr int DPP_SHL1 = 0x101;     // row_shl:1  lane l reads lane l+1 of its 16-lane row
constexpr int DPP_SHR1 = 0x111;     // row_shr:1  lane l reads lane l-1
constexpr int DPP_HMIRROR = 0x141;  // row_half_mirror  lane l reads lane 7-l of its 8-lane half
constexpr int QP_PAIR0 = 0xA0;      // quad_perm [0,0,2,2]

// sum over the G lanes of a group, in every lane of it
template <int G>
__device__ __forceinline__ double gsum(double p) {
  p += qperm<dpp16::QP_SWAP1>(p);
  if constexpr (G >= 4) p += qperm<dpp16::QP_SWAP2>(p);
  // after the quad sums, lane 7-l of an 8-lane half holds the other quad's sum
  if constexpr (G >= 8) p += qperm<DPP_HMIRROR>(p);
  return p;
}

// lane 0's value of a group, in every lane of it
template <int G>
__device__ __forceinline__ double gbcast0(double v, int gl) {
  if constexpr (G == 2) {
    return qperm<QP_PAIR0>(v);
  } else if constexpr (G == 4) {
    return qperm<dpp16::QP_LANE0>(v);
  } else {
    const double q = qperm<dpp16::QP_LANE0>(v);  // lane 0 of each quad
    const double m = qperm<DPP_HMIRROR>(q);      // lanes 4-7 read lanes 3-0
    return gl >= 4 ? m : q;
  }
}

__device__ __forceinline__ bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// ---------------------------------------------------------------------------
// Forward. Lane gl's window R holds y_u for u in [t-4-4gl, t-1-4gl] at slot
// u mod 4 (t0 is a multiple of 4, so step JJ of a ring pass has slot J = JJ & 3
// for y_t); c is lane gl's float4 of row t (lags 4gl+1 .. 4gl+4).
template <int J, int G>
__device__ __forceinline__ double fwd_step(double (&R)[4], const float4 c, float x, int gl) {
  // lag 4gl+1+m meets y_{t-1-4gl-m} in slot (J-1-m) & 3; the newest term goes last
  double p = (double)c.w * R[(J - 4) & 3];
  p = fma((double)c.z, R[(J - 3) & 3], p);
  p = fma((double)c.y, R[(J - 2) & 3], p);
  p = fma((double)c.x, R[(J - 1) & 3], p);
  const double y = (double)x - gsum<G>(p);
  // slot J (y_{t-4-4gl}) takes y_{t-4gl} from lane gl-1, or y_t in lane 0
  const double up = qperm<DPP_SHR1>(R[J]);
  R[J] = gl == 0 ? y : up;
  return y;
}

template <int NP, bool AL, int JJ>
__device__ __forceinline__ void fwd_ring(double (&R)[4], float4 (&cq)[Grp<NP>::RING], float (&xq)[Grp<NP>::RING],
                                         float& ylast, const float* Ab, const float* xb, float* yb, int64_t T,
                                         int64_t tb, int gl) {
  constexpr int RING = Grp<NP>::RING;
  if constexpr (JJ < RING) {
    constexpr int G = Grp<NP>::G;
    const int64_t t = tb + JJ;
    const double y = fwd_step<JJ & 3, G>(R, cq[JJ], xq[JJ], gl);
    const int64_t tn = min(t + RING, T - 1);
    cq[JJ] = ldq<AL>(Ab + tn * NP + 4 * gl);
    xq[JJ] = xb[tn];
    pin();
    // every lane of the group stores the same value; a step past the end stores
    // the last live output at T-1 again
    const bool live = t < T;
    ylast = live ? (float)y : ylast;
    yb[live ? t : T - 1] = ylast;
    fwd_ring<NP, AL, JJ + 1>(R, cq, xq, ylast, Ab, xb, yb, T, tb, gl);
  }
}

template <int NP, bool AL>
__device__ __forceinline__ void fwd_chunk(const float* Ab, const float* xb, float* yb, const double* cr,
                                          int64_t T, int64_t t0, int len, int gl) {
  constexpr int RING = Grp<NP>::RING;
  double R[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) R[3 - m] = cr[4 * gl + m];  // y_{t0-1-4gl-m}, slot 3-m
  float4 cq[RING];
  float xq[RING];
#pragma unroll
  for (int j = 0; j < RING; ++j) {
    const int64_t tc = min(t0 + j, T - 1);
    cq[j] = opq(ldq<AL>(Ab + tc * NP + 4 * gl));
    xq[j] = opq(xb[tc]);
  }
  settle(R);
  float ylast = 0.f;
  for (int jb = 0; jb < len; jb += RING) fwd_ring<NP, AL, 0>(R, cq, xq, ylast, Ab, xb, yb, T, t0 + jb, gl);
}

template <int NP>
__global__ __launch_bounds__(256) void grp_fwd_rerun(ChunkArgs<float> a) {
  constexpr int G = Grp<NP>::G;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t cid = gt / G;
  const int gl = (int)(gt - cid * G);
  if (cid >= (a.rg ? a.G : a.B * a.P)) return;  // a group's lanes leave together
  const int64_t b = chunk_signal(a, cid);
  const Span sp = sig_span(a, b);
  const int64_t t0 = (cid - sp.cbase) * a.L;
  const int64_t rem = sp.T - t0;
  const int len = rem < a.L ? (int)rem : a.L;
  const float* Ab = a.A + sp.ao;
  const double* cr = a.carry + cid * NP;
  if (aligned16(Ab))
    fwd_chunk<NP, true>(Ab, a.x + sp.xo, a.y + sp.xo, cr, sp.T, t0, len, gl);
  else
    fwd_chunk<NP, false>(Ab, a.x + sp.xo, a.y + sp.xo, cr, sp.T, t0, len, gl);
}

// ---------------------------------------------------------------------------
// Adjoint, descending, with two lags per lane: G = NP / 2 lanes per chunk,
// lane gl owning lags 2gl+1 and 2gl+2, so a group loads a row as G
// float2 and stores its dA row the same way. The adjoint needs no group sum,
// only lambda broadcast from the group's lane 0, so twice the forward's lanes
// per chunk cost one DPP move per step and give the latency-bound steps twice
// the waves. `now` in lane 0 is the pending sum of position t; lane gl's window
// P holds the pending sums of positions [t-2-2gl, t-1-2gl] at slot u mod 2, and
// (with DA) its window Y the y values of the same positions.
template <int NP>
struct AdjGrp {
  static constexpr int G = NP / 2;
  static constexpr int RING = 8;     // steps in flight, and whose outputs are written together
  // orders 16 and 32 stage their outputs; at order 8 the stage's writes cost more
  // issue slots than the light steps leave (order-8 adjoint of the ragged
  // workload: 117 us with a lane's own stores, 168 us staged)
  static constexpr bool STAGE = NP >= 16;
  static constexpr int CPW = 64 / G;  // chunks per wave (one wave per workgroup)
  // LDS staging of a ring turn's outputs (padded so that the chunks' rows fall
  // on different banks): dA rows and dx values, and per chunk the region to write
  static constexpr int SA = RING * NP + 2;
  static constexpr int SX = RING + 1;
};

// A ring turn's outputs are staged in LDS and then written by the whole wave,
// one chunk region after another, in 128-byte-aligned runs of 64 floats: each
// write covers whole cache lines except at a region's two ends. Written a row
// per chunk and step (32 to 128 bytes, at the packed rows' arbitrary alignment),
// the partial-line write requests had bounded the adjoint: a write-only probe of
// that pattern reached 2.0 TB/s with 32-byte rows, 3.2 TB/s with unaligned
// 64-byte rows, against 5.5 TB/s sequential.
template <int NP>
struct AdjStage {
  float dA[AdjGrp<NP>::CPW * AdjGrp<NP>::SA];
  float dx[AdjGrp<NP>::CPW * AdjGrp<NP>::SX];
  int64_t a0[AdjGrp<NP>::CPW];  // dA element of the region's first row
  int64_t x0[AdjGrp<NP>::CPW];  // dx element of the region's first row
  int nl[AdjGrp<NP>::CPW];      // rows of the region inside the signal
};

// The workgroup is one wave, whose LDS accesses complete in issue order: the
// staging only needs the compiler to keep them in program order. (A workgroup
// barrier would also wait for the ring's loads in flight.)
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint64_t u = (uint64_t)v;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Raw buffer over n floats at p: stores at byte offsets at or past n*4 are
// dropped by the hardware's range check, which masks lanes without a branch
// (a branch in the step loop's body lets the register allocator and the
// wait-count pass drain the prefetch ring).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t float_buffer(float* p, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, n * 4, 0x00020000);
}
constexpr int kDropped = (int)0x80000000;  // a byte offset past every buffer here
__device__ __forceinline__ void buffer_store(float v, __amdgpu_buffer_rsrc_t r, int byte_off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, byte_off, 0, 0);
}


template <int G>
__device__ __forceinline__ double gbcast0_adj(double v, int gl) {
  if constexpr (G == 16) {
    double r = 0.0;
    dpp16::fma_bcast<0>(r, v, 1.0);  // lane 0 of the 16-lane row, times one
    return r;
  } else {
    return gbcast0<G>(v, gl);
  }
}

template <bool AL>
__device__ __forceinline__ float2 ld2(const float* p) {
  if constexpr (AL) return *reinterpret_cast<const float2*>(p);
  return make_float2(p[0], p[1]);
}
template <bool AL>
__device__ __forceinline__ void st2(float* p, float2 v) {
  if constexpr (AL) {
    *reinterpret_cast<float2*>(p) = v;
  } else {
    p[0] = v.x;
    p[1] = v.y;
  }
}

struct AdjStreams {
  const float* Ab;
  const float* gb;
  const float* yb;
  const float* zib;
  const float* zb;  // zib, or yb when there is no zi (an always-valid address)
  int64_t dxo;      // element of dx / dA (in a.dx / a.dA) at the signal's sample 0
  int64_t dAo;
  float* dxb;       // the signal's dx and dA, for the stores of an unstaged group
  float* dAb;
  float* sink;      // the group's consumed carry slot: destination of the stores of steps past T
  int64_t T;
  int n;
  int cw;           // the chunk's slot in its wave
};

// DX / DA: whether dx / dA are written (template arguments, so that the step
// loop holds no branch); sa / sx: the LDS staging of step tb's outputs (the
// ring's later steps stage one row down each).
template <int NP, bool AL, bool DA, bool DX, int JJ>
__device__ __forceinline__ void adj_ring(double (&P)[2], double& now, float (&Y)[2],
                                         float2 (&cq)[AdjGrp<NP>::RING], float (&gq)[AdjGrp<NP>::RING],
                                         float (&yq)[AdjGrp<NP>::RING], const AdjStreams& io, int64_t tb,
                                         int gl, float* sa, float* sx) {
  constexpr int RING = AdjGrp<NP>::RING;
  if constexpr (JJ < RING) {
    constexpr int G = AdjGrp<NP>::G;
    constexpr int J = (1 - JJ) & 1;  // slot of t = tb - JJ (tb odd)
    const int64_t t = tb - JJ;
    const bool live = t < io.T;
    const double lam0 = gbcast0_adj<G>((double)gq[JJ] + now, gl);
    const double lam = live ? lam0 : 0.0;
    const float2 c = cq[JJ];
    P[(J - 1) & 1] = fma(-(double)c.x, lam, P[(J - 1) & 1]);
    P[J] = fma(-(double)c.y, lam, P[J]);
    // slot S (position t-1-2gl) leaves at the top: lane 0's becomes `now`, the
    // others take their upper neighbour's, the last lane a fresh zero
    constexpr int S = (J - 1) & 1;
    const double down = qperm<DPP_SHL1>(P[S]);
    now = P[S];
    P[S] = gl == G - 1 ? 0.0 : down;
    const float ynew = yw_flag(io.zib != nullptr, io.n, io.T, t - 1 - NP) ? yq[JJ] : 0.f;
    // refill the ring for step t - RING (times clamped into the signal; the y
    // value entering lane G-1's window at that step is y_{t-RING-1-NP})
    {
      const int64_t tn = max(min(t - RING, io.T - 1), (int64_t)0);
      cq[JJ] = ld2<AL>(io.Ab + tn * NP + 2 * gl);
      gq[JJ] = io.gb[tn];
      if constexpr (DA) yq[JJ] = *yw_ptr(io.yb, io.zb, io.n, io.T, t - RING - 1 - NP);
    }
    pin();
    const float lf = (float)lam;
    constexpr bool STAGE = AdjGrp<NP>::STAGE;
    if constexpr (DX) {
      if constexpr (STAGE)
        sx[-JJ] = lf;  // the group's lanes hold the same value
      else
        *(live ? io.dxb + t : io.sink) = lf;
    }
    if constexpr (DA) {
      const float2 d = make_float2(-lf * Y[(J - 1) & 1], -lf * Y[J]);
      if constexpr (STAGE)
        *reinterpret_cast<float2*>(sa - JJ * NP) = d;
      else
        st2<AL>(live ? io.dAb + t * NP + 2 * gl : io.sink + 2 * gl, d);
      // the window moves down one position: slot S takes y_{t-3-2gl} from the lane
      // above, the last lane y_{t-1-NP}
      const float ydown = qperm<DPP_SHL1>(Y[S]);
      Y[S] = gl == G - 1 ? ynew : ydown;
    }
    adj_ring<NP, AL, DA, DX, JJ + 1>(P, now, Y, cq, gq, yq, io, tb, gl, sa, sx);
  }
}

// Writes the ring turn's staged outputs (rows tb-RING+1 .. tb of every chunk of the
// wave that lie in its signal; `live`: the chunk exists).
template <int NP, bool DA, bool DX>
__device__ __forceinline__ void adj_flush(const AdjStreams& io, int64_t tb, bool live, AdjStage<NP>* stg, float* dA,
                                          float* dx) {
  using Q = AdjGrp<NP>;
  constexpr int RING = Q::RING;
  const int lane = (int)(threadIdx.x & 63);
  const int64_t r0 = tb - RING + 1;
  const int64_t nl64 = io.T - r0;
  const int nl = live ? (int)(nl64 < 0 ? 0 : nl64 > RING ? RING : nl64) : 0;
  // the group's lanes write the same values
  stg->a0[io.cw] = io.dAo + r0 * NP;
  stg->x0[io.cw] = io.dxo + r0;
  stg->nl[io.cw] = nl;
  lds_order();
#pragma unroll
  for (int c = 0; c < Q::CPW; ++c) {
    const int n = __builtin_amdgcn_readfirstlane(stg->nl[c]);
    if constexpr (DA) {
      // from the 128-byte line holding the region's first element
      const int64_t g0 = uniform64(stg->a0[c]);
      const int lo = (int)(g0 & 31);
      const __amdgpu_buffer_rsrc_t r = float_buffer(dA + (g0 - lo), lo + n * NP);
      constexpr int PIECES = (RING * NP + 31 + 63) / 64;
#pragma unroll
      for (int k = 0; k < PIECES; ++k) {
        const int i = k * 64 + lane;
        const int j = i - lo;
        const float v = stg->dA[c * Q::SA + min(max(j, 0), RING * NP - 1)];
        buffer_store(v, r, j >= 0 ? i * 4 : kDropped);
      }
    }
    if constexpr (DX) {
      const __amdgpu_buffer_rsrc_t r = float_buffer(dx + uniform64(stg->x0[c]), n);
      const int k = lane & (RING - 1);
      buffer_store(stg->dx[c * Q::SX + k], r, lane < RING ? k * 4 : kDropped);
    }
  }
  lds_order();
}

template <int NP, bool AL, bool DA, bool DX>
__device__ __forceinline__ void adj_chunk(const AdjStreams& io, const double* cr, int64_t t0, int L, int gl,
                                          float* dz, bool live, AdjStage<NP>* stg, float* dA, float* dx) {
  // tstart = t0 + L - 1 is odd: position tstart - (2gl+1+m) sits in slot m & 1
  // (m = 0, 1), in P as in Y
  double P[2];
  double now = cr[0];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int rr = 2 * gl + 1 + m;
    P[m] = rr < NP ? cr[rr] : 0.0;  // pending sum of position tstart-rr
  }
  const int64_t tstart = t0 + L - 1;
  float Y[2] = {0.f, 0.f};
  if constexpr (DA) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int64_t u = tstart - 1 - 2 * gl - m;
      const float v = *yw_ptr(io.yb, io.zb, io.n, io.T, u);
      Y[m] = yw_flag(io.zib != nullptr, io.n, io.T, u) ? v : 0.f;
    }
  }
  constexpr int RING = AdjGrp<NP>::RING;
  float2 cq[RING];
  float gq[RING], yq[RING];
#pragma unroll
  for (int j = 0; j < RING; ++j) {
    const int64_t tc = max(min(tstart - j, io.T - 1), (int64_t)0);
    cq[j] = opq(ld2<AL>(io.Ab + tc * NP + 2 * gl));
    gq[j] = opq(io.gb[tc]);
    yq[j] = 0.f;
    if constexpr (DA) yq[j] = opq(*yw_ptr(io.yb, io.zb, io.n, io.T, tstart - j - 1 - NP));
  }
  settle(P);
  if constexpr (AdjGrp<NP>::STAGE) {
    // staging of the turn's last row (rows ascend with t)
    float* const sa = &stg->dA[io.cw * AdjGrp<NP>::SA + (RING - 1) * NP + 2 * gl];
    float* const sx = &stg->dx[io.cw * AdjGrp<NP>::SX + RING - 1];
    for (int jb = 0; jb < L; jb += RING) {
      adj_ring<NP, AL, DA, DX, 0>(P, now, Y, cq, gq, yq, io, tstart - jb, gl, sa, sx);
      adj_flush<NP, DA, DX>(io, tstart - jb, live, stg, dA, dx);
    }
  } else {
    for (int jb = 0; jb < L; jb += RING)
      adj_ring<NP, AL, DA, DX, 0>(P, now, Y, cq, gq, yq, io, tstart - jb, gl, nullptr, nullptr);
  }
  if (dz && live) {
    // pending sums of positions -1-r: r = 0 in `now` (lane 0), 2gl+1+m in slot m
    if (gl == 0) dz[0] = (float)now;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int rr = 2 * gl + 1 + m;
      if (rr < io.n) dz[rr] = (float)P[m];
    }
  }
}

template <int NP>
__device__ __forceinline__ AdjStage<NP>* adj_stage() {
  __shared__ AdjStage<NP> stg;
  return &stg;
}

// One wave per workgroup (the staging is per wave). With staging every lane
// stays to the end, since the wave writes its chunks' outputs together: lanes
// past the last chunk run the last chunk again and write nothing.
template <int NP, bool DA>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void grp_bwd_rerun(ChunkArgs<float> a) {
  constexpr int G = AdjGrp<NP>::G;
  constexpr bool STAGE = AdjGrp<NP>::STAGE;
  const int64_t gt = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const int64_t nchunks = a.rg ? a.G : a.B * a.P;
  const int64_t cid0 = gt / G;
  const bool live = cid0 < nchunks;
  if (!STAGE && !live) return;  // a group's lanes leave together
  const int64_t cid = live ? cid0 : nchunks - 1;
  const int gl = (int)(gt - cid0 * G);
  const int64_t b = chunk_signal(a, cid);
  const Span sp = sig_span(a, b);
  const int64_t c = cid - sp.cbase;
  double* cr = a.carry + cid * NP;
  AdjStage<NP>* stg = nullptr;
  if constexpr (STAGE) stg = adj_stage<NP>();
  AdjStreams io;
  io.Ab = a.A + sp.ao;
  io.gb = a.x + sp.xo;  // upstream gradient
  io.yb = a.y + sp.xo;
  io.zib = a.zi ? a.zi + sp.zo : nullptr;
  io.zb = io.zib ? io.zib : io.yb;
  io.dxo = sp.xo;
  io.dAo = sp.dao;
  io.dxb = a.dx ? a.dx + sp.xo : nullptr;
  io.dAb = DA ? a.dA + sp.dao : nullptr;
  // the carry was read before any store: its NP doubles hold the group's 2G floats
  io.sink = reinterpret_cast<float*>(cr);
  io.T = sp.T;
  io.n = a.n;
  io.cw = (int)(threadIdx.x / G);
  float* dz = (c == 0 && a.dzi) ? a.dzi + sp.zo : nullptr;
  // Staged: one variant for the whole wave, whose lanes must not diverge, since
  // each flush writes every chunk of the wave from the region bounds its lanes
  // stage (a ragged wave's chunks belong to signals whose rows start at
  // different alignments, so the aligned loads are taken only when every chunk
  // allows them). Unstaged: the lane's own loads and dA stores.
  const bool al =
      STAGE ? __builtin_amdgcn_ballot_w64((reinterpret_cast<uintptr_t>(io.Ab) & 7) != 0) == 0
            : ((reinterpret_cast<uintptr_t>(io.Ab) | (DA ? reinterpret_cast<uintptr_t>(io.dAb) : 0)) & 7) == 0;
  if (a.dx) {
    if (al)
      adj_chunk<NP, true, DA, true>(io, cr, c * a.L, a.L, gl, dz, live, stg, a.dA, a.dx);
    else
      adj_chunk<NP, false, DA, true>(io, cr, c * a.L, a.L, gl, dz, live, stg, a.dA, a.dx);
  } else {
    if (al)
      adj_chunk<NP, true, DA, false>(io, cr, c * a.L, a.L, gl, dz, live, stg, a.dA, a.dx);
    else
      adj_chunk<NP, false, DA, false>(io, cr, c * a.L, a.L, gl, dz, live, stg, a.dA, a.dx);
  }
}

// Adjoint zero-state responses of chunks 1 .. P-1 for a backward that has the
// forward's transition matrices (the adjoint carry chain reads them transposed):
// the group's adjoint recurrence over the chunk from zero pending sums, driven
// by g, leaves w[r], the pending sum of position t0-1-r, which is stored in the
// chunk's summary slot after its matrix (where the forward kept z). N FMAs per
// sample instead of the N^2 of a summary that also forms the matrix.
template <int NP, bool AL>
__device__ __forceinline__ void wsum_chunk(const AdjStreams& io, int64_t t0, int L, int gl, double* out) {
  constexpr int RING = AdjGrp<NP>::RING;
  double P[2] = {0.0, 0.0};
  double now = 0.0;
  float Y[2] = {0.f, 0.f};
  const int64_t tstart = t0 + L - 1;
  float2 cq[RING];
  float gq[RING], yq[RING];
#pragma unroll
  for (int j = 0; j < RING; ++j) {
    const int64_t tc = max(min(tstart - j, io.T - 1), (int64_t)0);
    cq[j] = opq(ld2<AL>(io.Ab + tc * NP + 2 * gl));
    gq[j] = opq(io.gb[tc]);
    yq[j] = 0.f;
  }
  for (int jb = 0; jb < L; jb += RING)
    adj_ring<NP, AL, false, false, 0>(P, now, Y, cq, gq, yq, io, tstart - jb, gl, nullptr, nullptr);
  // pending sums of positions t0-1-r: r = 0 in `now` (lane 0), 2gl+1+m in slot m
  if (gl == 0) out[0] = now;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int rr = 2 * gl + 1 + m;
    if (rr < NP) out[rr] = P[m];
  }
}

template <int NP>
__global__ __launch_bounds__(256) void grp_wsum(ChunkArgs<float> a) {
  constexpr int G = AdjGrp<NP>::G;
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t cid = gt / G;
  const int gl = (int)(gt - cid * G);
  if (cid >= a.G) return;  // a group's lanes leave together
  const int64_t b = chunk_signal(a, cid);
  const Span sp = sig_span(a, b);
  const int64_t c = cid - sp.cbase;
  if (c < 1) return;  // the chain stops at chunk 0's carry
  AdjStreams io{};
  io.Ab = a.A + sp.ao;
  io.gb = a.x + sp.xo;  // upstream gradient
  io.T = sp.T;
  io.n = a.n;
  double* out = a.summ + summ_slot(a, b, sp, c, false) * (NP * NP + NP) + NP * NP;
  if ((reinterpret_cast<uintptr_t>(io.Ab) & 7) == 0)
    wsum_chunk<NP, true>(io, c * a.L, a.L, gl, out);
  else
    wsum_chunk<NP, false>(io, c * a.L, a.L, gl, out);
}

inline dim3 grp_grid(const ChunkArgs<float>& a, int G, unsigned threads) {
  const int64_t lanes = (a.rg ? a.G : a.B * a.P) * G;
  return dim3((unsigned)((lanes + threads - 1) / threads));
}
// 64-lane workgroups: the launches hold some 32K lanes, and 256-lane ones would
// leave a few hundred workgroups to balance over 256 CUs (order 32 in the ragged
// workload, 384-sample chunks: 1.11 ms per step against 0.88 ms at 448 samples,
// the first length under the former 64K-lane switch to 64-lane workgroups)
inline unsigned grp_threads(const ChunkArgs<float>&, int) { return 64u; }

}  // namespace grp

bool grp_rerun_ok(const ChunkArgs<float>& a, int np) {
  return (np == 2 || np == 4 || np == 8 || np == 16 || np == 32) && a.n == np && !a.neg && a.dapart == nullptr &&
         a.rx == nullptr && (a.rg != nullptr || a.a_tstride == a.n) && a.L % 16 == 0;
}

// orders 2 and 4: one quad per chunk (tlpc_quad_rerun.hip)
int launch_quad_fwd_rerun(ChunkArgs<float> a, int np, hipStream_t st);
int launch_quad_bwd_rerun(ChunkArgs<float> a, int np, hipStream_t st);

int launch_grp_fwd_rerun(ChunkArgs<float> a, int np, hipStream_t st) {
  if (np == 2 || np == 4) return launch_quad_fwd_rerun(a, np, st);
  const int G = np / 4;
  const unsigned th = grp::grp_threads(a, G);
  const dim3 g = grp::grp_grid(a, G, th);
  if (g.x == 0) return 0;
  switch (np) {
    case 8: hipLaunchKernelGGL(grp::grp_fwd_rerun<8>, g, dim3(th), 0, st, a); break;
    case 16: hipLaunchKernelGGL(grp::grp_fwd_rerun<16>, g, dim3(th), 0, st, a); break;
    case 32: hipLaunchKernelGGL(grp::grp_fwd_rerun<32>, g, dim3(th), 0, st, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

bool grp_wsum_ok(const ChunkArgs<float>& a, int np) {
  return (np == 8 || np == 16 || np == 32) && a.rg != nullptr && grp_rerun_ok(a, np);
}

int launch_grp_wsum(ChunkArgs<float> a, int np, hipStream_t st) {
  const int G = np / 2;
  const unsigned th = grp::grp_threads(a, G);
  const dim3 g = grp::grp_grid(a, G, th);
  if (g.x == 0) return 0;
  switch (np) {
    case 8: hipLaunchKernelGGL(grp::grp_wsum<8>, g, dim3(th), 0, st, a); break;
    case 16: hipLaunchKernelGGL(grp::grp_wsum<16>, g, dim3(th), 0, st, a); break;
    case 32: hipLaunchKernelGGL(grp::grp_wsum<32>, g, dim3(th), 0, st, a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int launch_grp_bwd_rerun(ChunkArgs<float> a, int np, hipStream_t st) {
  if (np == 2 || np == 4) return launch_quad_bwd_rerun(a, np, st);
  const int G = np / 2;
  const unsigned th = grp::grp_threads(a, G);
  const dim3 g = grp::grp_grid(a, G, th);
  if (g.x == 0) return 0;
  const bool da = a.dA != nullptr;
  switch (np) {
    case 8:
      if (da) hipLaunchKernelGGL((grp::grp_bwd_rerun<8, true>), g, dim3(th), 0, st, a);
      else hipLaunchKernelGGL((grp::grp_bwd_rerun<8, false>), g, dim3(th), 0, st, a);
      break;
    case 16:
      if (da) hipLaunchKernelGGL((grp::grp_bwd_rerun<16, true>), g, dim3(th), 0, st, a);
      else hipLaunchKernelGGL((grp::grp_bwd_rerun<16, false>), g, dim3(th), 0, st, a);
      break;
    case 32:
      if (da) hipLaunchKernelGGL((grp::grp_bwd_rerun<32, true>), g, dim3(th), 0, st, a);
      else hipLaunchKernelGGL((grp::grp_bwd_rerun<32, false>), g, dim3(th), 0, st, a);
      break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // namespace tlpc
