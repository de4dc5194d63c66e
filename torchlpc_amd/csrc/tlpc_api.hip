// C ABI of libtorchlpc_amd_hip.so (declared in include/torchlpc_amd.h).
// Validates sizes, carves the caller's workspace, picks a strategy and launches.
#include <stddef.h>

#include <atomic>
#include <cmath>
#include <mutex>
#include <stdint.h>

#include "../../include/torchlpc_amd.h"
#include "tlpc_kernels.h"

// Workspace of tlpc_lpc_forward_noalloc: one copy per GPU, held until the
// process ends. In host code `scratch` is only a stand-in for the symbol; the
// device address comes from hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) unsigned char scratch[TLPC_NOALLOC_SCRATCH_BYTES];

namespace tlpc {
namespace {

constexpr size_t kAlign = 256;
constexpr int kRecomputeLmax = 256;  // longest chunk of the y-rebuilding adjoint rerun
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

inline int compiled_order(int n) {
  if (n <= 1) return 1;
  if (n <= 2) return 2;
  if (n <= 4) return 4;
  if (n <= 8) return 8;
  if (n <= 16) return 16;
  if (n <= 32) return 32;
  return 0;  // served by the serial strategy
}

// element size of the summaries and carries (float data keeps them in double)
inline size_t wide_size(int dtype) {
  switch (dtype) {
    case TLPC_F32: return 8;
    case TLPC_F64: return 8;
    case TLPC_C64: return 16;
    case TLPC_C128: return 16;
    default: return 0;
  }
}

inline size_t elem_size(int dtype) {
  switch (dtype) {
    case TLPC_F32: return 4;
    case TLPC_F64: return 8;
    case TLPC_C64: return 8;
    case TLPC_C128: return 16;
    default: return 0;
  }
}

// Chunk length: a multiple of the staging block, at least 4*NP and at most 4096
// samples, sized for a number of chunks: 32K with one rerun lane per chunk, and
// for the float32 multi-lane kernels (orders 2 to 32 that fill their padded
// width, n == NP) 16K at orders 2 and 4, 8K at orders 8 and 16 and 4K at order
// 32, where longer chunks shorten the carry chains while the lanes per chunk
// keep the reruns busy. These are the fastest lengths measured on the ragged
// workload's order groups (about 3.3M samples each): 240 / 224 / 416 / 400 /
// 832 samples at orders 2 / 4 / 8 / 16 / 32, 0.087 / 0.111 / 0.249 / 0.361 /
// 0.639 ms per group step, against 0.095 / 0.131 / 0.37 / 0.61 / 1.06 ms at
// 112-128 samples. Summaries and carries are in the wide type (float64 for
// single-precision data), so a chunk boundary adds no single-precision rounding
// and shorter chunks cost no accuracy.
inline bool multi_lane_rerun(int dtype, int n, int np) { return dtype == TLPC_F32 && n == np && np >= 2; }
inline int pick_chunk(int64_t B, int64_t T, int np, int requested, bool multi_lane) {
  const int blk = chunk_block(np);
  const int64_t target = !multi_lane ? 32768 : np <= 4 ? 16384 : np <= 16 ? 8192 : 4096;  // chunks
  int64_t L = requested > 0 ? requested : (B * T + target - 1) / target;
  if (requested <= 0 && L < 4 * np) L = 4 * np;
  if (L < blk) L = blk;
  if (L > 4096 && requested <= 0) L = 4096;
  if (L > (int64_t)1 << 24) L = (int64_t)1 << 24;
  L = (L + blk - 1) / blk * blk;
  return (int)L;
}

// Compute units of the current device (cached per device).
inline int device_cus() {
  // one slot per device, filled on first use; concurrent callers store the same
  // value, so relaxed atomics suffice
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cus[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cus[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// Order-16 float32 chunks (the quad kernels): at most 64 chunks per compute
// unit, i.e. one 256-thread rerun workgroup per CU and one round of summary
// workgroups. These are fewer, longer chunks than the generic rule gives: the
// reruns stream A and dA for every chunk at once and fewer streams use the memory
// better, and the carry chains are shorter. The rerun time follows the busiest
// CU (a CU given two workgroups takes twice as long), so the chunks per signal
// are chosen first, P = 64 * CUs / B, and the length rounded up to 16 samples
// from there (which can only lower P). At least 16 * 16 samples.
inline int pick_chunk16(int64_t B, int64_t T, int requested) {
  if (requested > 0) return pick_chunk(B, T, 16, requested, true);
  int64_t P = 64 * (int64_t)device_cus() / (B > 0 ? B : 1);
  if (P < 1) P = 1;
  int64_t L = (T + P - 1) / P;
  if (L < 256) L = 256;
  if (L > 4096) L = 4096;
  return (int)((L + 15) / 16 * 16);
}

// Time-invariant chunks: the summary cost no longer grows with the order (Phi is
// built once per signal), so chunks are short enough for about 48K lanes in the
// zero-state and rerun phases, at least 4*NP long, and at most 256 per signal
// (the length of the carry chain). More lanes hide more of the reruns' latency
// but lengthen the carry chains: at B=1024, T=16384, N=32 a step took 0.475 ms
// with chunks of 128 samples (128K lanes), 0.357 / 0.350 ms at 320 / 384 and
// 0.388 ms at 512. Summaries and carries are in the wide type,
// so short float32 chunks keep their accuracy.
// Time-invariant filters: about 49152 chunk lanes over the batch, and at least
// sqrt(1.5 T) samples per chunk. A signal's time is about alpha * T / L (its
// serial chain of T / L carry links) plus beta * L (a lane's serial steps through
// one chunk, zero-state pass and rerun, both directions); measured on long signals
// (B = 4, T = 2^20, orders 2 and 32) the fastest chunk length sits near
// sqrt(1.5 T), with a flat optimum between 0.7x and 1.2x of it.
inline int pick_chunk_ti(int64_t B, int64_t T, int np, int requested) {
  const int blk = chunk_block(np);
  int64_t L = requested > 0 ? requested : (B * T + 49151) / 49152;
  if (requested <= 0) {
    if (L < 4 * np) L = 4 * np;
    const int64_t Lmin = (int64_t)std::ceil(std::sqrt(1.5 * (double)T));
    if (L < Lmin) L = Lmin;
  }
  if (L < blk) L = blk;
  if (L > (int64_t)1 << 24) L = (int64_t)1 << 24;
  return (int)((L + blk - 1) / blk * blk);
}

// The per-signal kernels (one workgroup per signal, no workspace) serve float32
// filters: order-16 time-varying ones (tiled, every sample of A read once per
// direction) and time-invariant ones of order 16..32 (block products on MFMA;
// the reduced da only, not a per-sample dA).
inline bool signal_ok(int dtype, int n, int flags, int neg) {
  if (dtype != TLPC_F32 || neg) return false;
  if (flags & TLPC_TIME_INVARIANT) return n >= 16 && n <= 32 && !(flags & TLPC_DA_FULL);
  return n == 16;
}

// AUTO takes the per-signal kernels when the signals fill the chip: one
// workgroup (one signal) per compute unit, so a batch of B signals takes
// ceil(B / CUs) rounds of the per-signal time. Measured at T = 48000 the chunked
// kernels take about 1.8 times the per-signal time for a batch of CUs signals
// (and scale with B), so the per-signal kernels win while
// ceil(B / CUs) < 1.8 * B / CUs.
inline bool signal_auto(int64_t B) {
  const int64_t cus = device_cus();
  if (B < cus) return false;
  const int64_t rounds = (B + cus - 1) / cus;
  return rounds * cus * 10 < 18 * B;
}

inline int pick_strategy(int strategy, int np, bool sig_ok, int64_t B) {
  if (np == 0) return TLPC_STRATEGY_SERIAL;
  if (strategy == TLPC_STRATEGY_SERIAL) return TLPC_STRATEGY_SERIAL;
  if (strategy == TLPC_STRATEGY_SIGNAL && sig_ok) return TLPC_STRATEGY_SIGNAL;
  if (strategy == TLPC_STRATEGY_AUTO && sig_ok && signal_auto(B)) return TLPC_STRATEGY_SIGNAL;
  return TLPC_STRATEGY_CHUNKED;
}

struct Plan {
  int np = 0;
  int strategy = 0;
  int L = 0;
  int64_t P = 0;
  bool c16 = false;  // float32 order-16 time-varying kernels
  size_t summ_bytes = 0;
  size_t summ_off = 0, carry_off = 0, wsum_off = 0, dapart_off = 0, ypad_off = 0, segs_off = 0,
         total = 0;
  bool ypad = false;  // time-invariant adjoint with da partials: the y window's zi blocks
  int64_t nseg = 0;   // segments per carry chain (0: one workgroup per chain)
};

// Segments of a long carry chain (time-varying, uniform batch, orders up to 4,
// which take the blocked carry scan): about 512 links each, at least 4 and at
// most 256 of them; 0 keeps one workgroup per chain. The segmented launches put
// the signals on the grid's y axis (at most 65535).
inline int64_t carry_segments(int np, int64_t B, int64_t P, bool ti, bool c16) {
  if (ti || c16 || np > 4 || B > 65535) return 0;
  const int64_t links = P - 1;
  const int64_t nseg = (links + 511) / 512;
  if (nseg < 4) return 0;
  return nseg > 256 ? 256 : nseg;
}

// The order-16 float32 kernels serve time-varying filters with chunks that are
// whole 16-step blocks and 32-bit sample indices inside a signal.
inline bool chunk16_ok(int dtype, int32_t n, int32_t flags, int neg, int L, int64_t T) {
  return dtype == TLPC_F32 && n == 16 && !(flags & TLPC_TIME_INVARIANT) && !neg && L % 16 == 0 &&
         T * 16 < (int64_t(1) << 31);
}

Plan make_plan(int dtype, int64_t B, int64_t T, int32_t n, int32_t flags, int32_t strategy,
               int32_t chunk_len, bool backward, int neg = 0) {
  Plan p;
  const size_t es = elem_size(dtype);
  p.np = compiled_order(n);
  p.strategy = pick_strategy(strategy, p.np, signal_ok(dtype, n, flags, neg), B);
  if (p.strategy == TLPC_STRATEGY_SIGNAL) {
    p.P = 1;
    return p;  // summaries and carries stay on chip: no workspace
  }
  if (p.strategy == TLPC_STRATEGY_SERIAL) {
    // the serial adjoint keeps lambda in the workspace when dx is not requested
    if (backward && !(flags & TLPC_NEED_DX)) p.total = align_up((size_t)B * T * es);
    p.carry_off = 0;
    return p;
  }
  const bool ti = (flags & TLPC_TIME_INVARIANT) != 0;
  if (ti)
    p.L = pick_chunk_ti(B, T, p.np, chunk_len);
  else if (chunk16_ok(dtype, n, flags, neg, 256, T))
    p.L = pick_chunk16(B, T, chunk_len);
  else {
    const bool ml = multi_lane_rerun(dtype, n, p.np);
    p.L = pick_chunk(B, T, p.np, chunk_len, ml);
    // A small uniform batch of long signals would get many short chunks and
    // long carry chains: at least min(64, T / 128) samples per chunk there
    // (B=16, T=8192, N=4 forward: 512 chunks of 16 samples 29.6 us, 128 of 64
    // samples 25.1 us, one HIP graph each)
    if (ml && chunk_len <= 0) {
      const int blk = chunk_block(p.np);
      const int64_t lmin = std::min<int64_t>(64, T / 128);
      if (p.L < lmin) p.L = (int)((lmin + blk - 1) / blk * blk);
    }
  }
  p.P = T > 0 ? (T + p.L - 1) / p.L : 0;
  const size_t ws_es = wide_size(dtype);
  // time-invariant summaries are the zero-state vectors only (Phi is per signal)
  const size_t summ = (size_t)B * (p.P > 1 ? p.P - 1 : 0) * (ti ? p.np : p.np * p.np + p.np) * ws_es;
  const size_t carry = (size_t)B * p.P * p.np * ws_es;
  const size_t dapart =
      (backward && (flags & TLPC_TIME_INVARIANT) && (flags & TLPC_NEED_DA) && !(flags & TLPC_DA_FULL))
          ? (size_t)B * p.P * p.np * ws_es
          : 0;
  p.c16 = p.strategy == TLPC_STRATEGY_CHUNKED && chunk16_ok(dtype, n, flags, neg, p.L, T);
  // the order-16 adjoint keeps its zero-state responses apart from the summaries,
  // which hold the forward's transition matrices
  const size_t wsum = (backward && p.c16) ? (size_t)B * (p.P > 1 ? p.P - 1 : 0) * p.np * ws_es : 0;
  p.summ_off = 0;
  p.summ_bytes = align_up(summ);
  p.carry_off = p.summ_bytes;
  p.wsum_off = p.carry_off + align_up(carry);
  p.dapart_off = p.wsum_off + align_up(wsum);
  p.ypad = dapart > 0;
  p.ypad_off = p.dapart_off + align_up(dapart);
  p.segs_off = p.ypad_off + (p.ypad ? align_up((size_t)B * ti_ypad(p.np) * es) : 0);
  p.nseg = carry_segments(p.np, B, p.P, ti, p.c16);
  p.total = p.segs_off + align_up((size_t)B * p.nseg * (p.np * p.np + 2 * p.np) * ws_es);
  return p;
}

// Ragged workspace: summaries and carries indexed by chunk id.
inline size_t ragged_workspace(size_t wsz, int64_t G, int np) {
  return align_up((size_t)G * (np * np + np) * wsz) + align_up((size_t)G * np * wsz);
}

// Device address of `scratch` on the current device, read once per device (with
// that device current) and kept; later calls return the kept address.
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_dev[kMaxDevices];

int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return TLPC_EDEVICE;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TLPC_EDEVICE;
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return TLPC_OK;
}

int check_common(int dtype, int64_t B, int64_t T, int32_t n) {
  if (elem_size(dtype) == 0) return TLPC_EDTYPE;
  if (B < 0 || T < 0 || n < 1) return TLPC_EINVAL;
  if (n > 4096) return TLPC_EORDER;
  return TLPC_OK;
}

template <typename S, int NP>
int run_chunked(const ChunkArgs<S>& a, bool backward, hipStream_t st) {
  return backward ? launch_chunked_bwd<S, NP>(a, st) : launch_chunked_fwd<S, NP>(a, st);
}

template <typename S>
int dispatch_np(const Plan& p, const ChunkArgs<S>& a, bool backward, hipStream_t st) {
  switch (p.np) {
    case 1: return run_chunked<S, 1>(a, backward, st);
    case 2: return run_chunked<S, 2>(a, backward, st);
    case 4: return run_chunked<S, 4>(a, backward, st);
    case 8: return run_chunked<S, 8>(a, backward, st);
    case 16: return run_chunked<S, 16>(a, backward, st);
    case 32: return run_chunked<S, 32>(a, backward, st);
    default: return TLPC_EORDER;
  }
}

// Adjoint without a saved y: the forward's input and its whole workspace, whose
// chunk carries the adjoint rerun rebuilds y from.
struct Recompute {
  const void* x;
  const void* fws;
  size_t fws_bytes;
};

// Whether the backward with these arguments can rebuild y from the forward's
// carries: the chunked scan over a time-varying filter of order up to 4 in real
// data, with chunks the rebuilding rerun takes, planned like the forward.
inline bool recompute_ok(int dtype, int64_t B, int64_t T, int32_t n, int32_t flags, int32_t strategy,
                         int32_t chunk_len) {
  if (dtype != TLPC_F32 && dtype != TLPC_F64) return false;
  if (flags & TLPC_TIME_INVARIANT) return false;
  if (B < 1 || T < 1 || n < 1 || n > 4) return false;
  const Plan b = make_plan(dtype, B, T, n, flags, strategy, chunk_len, true);
  const Plan f = make_plan(dtype, B, T, n, 0, strategy, chunk_len, false);
  return b.strategy == TLPC_STRATEGY_CHUNKED && f.strategy == TLPC_STRATEGY_CHUNKED && b.L == f.L &&
         b.P == f.P && b.L <= kRecomputeLmax && b.L % 16 == 0;
}

template <typename S>
int run(int dtype, const void* x, const void* A, const void* zi, void* y, void* dx, void* dA,
        void* dzi, int64_t B, int64_t T, int32_t n, int32_t flags, int32_t strategy,
        int32_t chunk_len, void* ws, size_t ws_bytes, void* stream, bool backward, int neg,
        const Recompute* rc = nullptr) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const Plan p = make_plan(dtype, B, T, n, flags, strategy, chunk_len, backward, neg);
  if (B == 0) return TLPC_OK;
  size_t fcarry_off = 0;
  if (rc) {
    if (!backward || !rc->x || !recompute_ok(dtype, B, T, n, flags, strategy, chunk_len)) return TLPC_EINVAL;
    const Plan f = make_plan(dtype, B, T, n, 0, strategy, chunk_len, false);
    if (!rc->fws || rc->fws_bytes < f.total) return TLPC_EWORKSPACE;
    fcarry_off = f.carry_off;
  }
  const bool ti = (flags & TLPC_TIME_INVARIANT) != 0;
  const bool da_reduced = ti && !(flags & TLPC_DA_FULL);
  const bool need_dx = backward && (flags & TLPC_NEED_DX);
  const bool need_da = backward && (flags & TLPC_NEED_DA);
  const bool need_dzi = backward && (flags & TLPC_NEED_DZI);
  // per-sample buffers of an empty signal hold no elements (torch gives them no
  // storage); only zi, dzi and a reduced da are per signal
  const bool none = T == 0;
  if ((!A && !(none && !ti)) || (!x && !none)) return TLPC_EINVAL;
  if (!backward && !y && !none) return TLPC_EINVAL;
  if (backward && need_da && !y && !rc && !none) return TLPC_EINVAL;
  if ((need_dx && !dx && !none) || (need_da && !dA && !(none && !da_reduced)) ||
      (need_dzi && !dzi))
    return TLPC_EINVAL;
  if (ws_bytes < p.total || (p.total > 0 && !ws)) return TLPC_EWORKSPACE;

  if (T == 0) {
    // nothing flows backwards through an empty signal: dzi and da are zero
    int rc = TLPC_OK;
    if (need_dzi) rc = launch_fill_zero<S>(static_cast<S*>(dzi), B * n, st);
    if (rc == TLPC_OK && need_da && da_reduced) rc = launch_fill_zero<S>(static_cast<S*>(dA), B * n, st);
    return rc;
  }

  ChunkArgs<S> a{};
  a.x = static_cast<const S*>(x);
  a.A = static_cast<const S*>(A);
  a.zi = static_cast<const S*>(zi);
  a.y = static_cast<S*>(y);
  a.dx = need_dx ? static_cast<S*>(dx) : nullptr;
  a.dzi = need_dzi ? static_cast<S*>(dzi) : nullptr;
  a.B = B;
  a.T = T;
  a.n = n;
  a.L = p.L;
  a.P = p.P;
  a.a_bstride = ti ? n : T * n;
  a.a_tstride = ti ? 0 : n;
  a.neg = neg;
  char* w = static_cast<char*>(ws);
  if (p.strategy == TLPC_STRATEGY_SERIAL) {
    a.dA = (need_da && !da_reduced) ? static_cast<S*>(dA) : nullptr;
    a.da_ti = (need_da && da_reduced) ? static_cast<S*>(dA) : nullptr;
    a.carry = reinterpret_cast<typename Wide<S>::type*>(w);
    return backward ? launch_serial_bwd<S>(a, st) : launch_serial_fwd<S>(a, st);
  }
  if (p.strategy == TLPC_STRATEGY_SIGNAL) {
    if constexpr (sizeof(S) == 4) {
      if (ti) {
        a.da_ti = need_da ? static_cast<S*>(dA) : nullptr;
        return backward ? launch_ti_mfma_bwd(a, st) : launch_ti_mfma_fwd(a, st);
      }
      a.dA = need_da ? static_cast<S*>(dA) : nullptr;
      return backward ? launch_tile16_bwd(a, st) : launch_tile16_fwd(a, st);
    }
    return TLPC_EINVAL;
  }
  using W = typename Wide<S>::type;
  if (rc) {
    a.rx = static_cast<const S*>(rc->x);
    a.fcarry = reinterpret_cast<const W*>(static_cast<const char*>(rc->fws) + fcarry_off);
    a.y = nullptr;
  }
  a.summ = reinterpret_cast<W*>(w + p.summ_off);
  a.carry = reinterpret_cast<W*>(w + p.carry_off);
  a.wsum = reinterpret_cast<W*>(w + p.wsum_off);
  if (p.nseg > 0) {
    a.segs = reinterpret_cast<W*>(w + p.segs_off);
    a.nseg = p.nseg;
  }
  if (need_da && da_reduced) {
    a.dapart = reinterpret_cast<W*>(w + p.dapart_off);
    if (p.ypad) a.ypad = reinterpret_cast<S*>(w + p.ypad_off);
    a.da_ti = static_cast<S*>(dA);
  } else if (need_da) {
    a.dA = static_cast<S*>(dA);
  }
  if constexpr (sizeof(S) == 4) {
    // float32, order 16, time-varying: DPP summaries and quad reruns. They read
    // A (and write dA) as one float4 per lane, so both must be 16-byte aligned;
    // a view that starts mid-row takes the generic kernels (same plan and
    // workspace, which hold summaries of the same size).
    const bool rows_aligned = (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
                              (reinterpret_cast<uintptr_t>(a.dA) & 15) == 0;
    if (p.c16 && rows_aligned) {
      int rc = TLPC_OK;
      if (!backward) {
        rc = launch_chunk16_fwd_summary(a, true, st);
      } else {
        if (!(flags & TLPC_REUSE_SUMMARY)) rc = launch_chunk16_fwd_summary(a, false, st);
        if (rc == TLPC_OK) rc = launch_chunk16_adj_summary(a, st);
      }
      if (rc == TLPC_OK) rc = launch_chunk16_carry(a, !backward, st);
      if (rc == TLPC_OK) rc = backward ? launch_chunk16_adj_rerun(a, st) : launch_chunk16_fwd_rerun(a, st);
      return rc;
    }
  }
  return dispatch_np<S>(p, a, backward, st);
}

// Ragged batch of one order: the generic chunked kernels driven by the device
// table `rg` (layout in tlpc_kernels.h). Time-varying coefficients only.
template <typename S>
int run_ragged(const void* x, const void* A, const void* zi, void* y, void* dx, void* dA,
               void* dzi, const int64_t* rg, int64_t K, int64_t G, int64_t maxP, int32_t n,
               int32_t flags, int32_t L, void* ws, size_t ws_bytes, void* stream, bool backward) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int np = compiled_order(n);
  if (np == 0) return TLPC_EORDER;
  if (K < 0 || G < 0 || maxP < 0 || L <= 0 || L % chunk_block(np) != 0) return TLPC_EINVAL;
  if (K == 0) return TLPC_OK;
  const size_t need = ragged_workspace(sizeof(typename Wide<S>::type), G, np);
  if (ws_bytes < need || !ws) return TLPC_EWORKSPACE;
  const bool need_dx = backward && (flags & TLPC_NEED_DX);
  const bool need_da = backward && (flags & TLPC_NEED_DA);
  const bool need_dzi = backward && (flags & TLPC_NEED_DZI);
  if (!rg || !A || !x || (!backward && !y) || (need_da && !y)) return TLPC_EINVAL;
  if ((need_dx && !dx) || (need_da && !dA) || (need_dzi && !dzi)) return TLPC_EINVAL;
  ChunkArgs<S> a{};
  a.x = static_cast<const S*>(x);
  a.A = static_cast<const S*>(A);
  a.zi = static_cast<const S*>(zi);
  a.y = static_cast<S*>(y);
  a.dx = need_dx ? static_cast<S*>(dx) : nullptr;
  a.dA = need_da ? static_cast<S*>(dA) : nullptr;
  a.dzi = need_dzi ? static_cast<S*>(dzi) : nullptr;
  a.B = K;
  a.T = 0;
  a.n = n;
  a.L = L;
  a.P = maxP;
  a.a_bstride = 0;
  a.a_tstride = n;
  a.rg = rg;
  a.G = G;
  a.kept = backward && (flags & TLPC_REUSE_SUMMARY) ? 1 : 0;
  using W = typename Wide<S>::type;
  char* w = static_cast<char*>(ws);
  a.summ = reinterpret_cast<W*>(w);
  a.carry = reinterpret_cast<W*>(w + align_up((size_t)G * (np * np + np) * sizeof(W)));
  if (G == 0) return TLPC_OK;
  return dispatch_np<S>(Plan{np, TLPC_STRATEGY_CHUNKED, L, maxP}, a, backward, st);
}

int run_ragged_dtype(int dtype, const void* x, const void* A, const void* zi, void* y, void* dx,
                     void* dA, void* dzi, const int64_t* rg, int64_t K, int64_t G, int64_t maxP,
                     int32_t n, int32_t flags, int32_t L, void* ws, size_t ws_bytes, void* stream,
                     bool backward) {
  const int rc = check_common(dtype, K, 1, n);
  if (rc != TLPC_OK) return rc;
  switch (dtype) {
    case TLPC_F32:
      return run_ragged<float>(x, A, zi, y, dx, dA, dzi, rg, K, G, maxP, n, flags, L, ws, ws_bytes, stream, backward);
    case TLPC_F64:
      return run_ragged<double>(x, A, zi, y, dx, dA, dzi, rg, K, G, maxP, n, flags, L, ws, ws_bytes, stream, backward);
    case TLPC_C64:
      return run_ragged<Cx<float>>(x, A, zi, y, dx, dA, dzi, rg, K, G, maxP, n, flags, L, ws, ws_bytes, stream, backward);
    case TLPC_C128:
      return run_ragged<Cx<double>>(x, A, zi, y, dx, dA, dzi, rg, K, G, maxP, n, flags, L, ws, ws_bytes, stream, backward);
    default:
      return TLPC_EDTYPE;
  }
}

int run_dtype(int dtype, const void* x, const void* A, const void* zi, void* y, void* dx,
              void* dA, void* dzi, int64_t B, int64_t T, int32_t n, int32_t flags,
              int32_t strategy, int32_t chunk_len, void* ws, size_t ws_bytes, void* stream,
              bool backward, int neg = 0) {
  const int rc = check_common(dtype, B, T, n);
  if (rc != TLPC_OK) return rc;
  switch (dtype) {
    case TLPC_F32:
      return run<float>(dtype, x, A, zi, y, dx, dA, dzi, B, T, n, flags, strategy, chunk_len, ws,
                        ws_bytes, stream, backward, neg);
    case TLPC_F64:
      return run<double>(dtype, x, A, zi, y, dx, dA, dzi, B, T, n, flags, strategy, chunk_len, ws,
                         ws_bytes, stream, backward, neg);
    case TLPC_C64:
      return run<Cx<float>>(dtype, x, A, zi, y, dx, dA, dzi, B, T, n, flags, strategy, chunk_len,
                            ws, ws_bytes, stream, backward, neg);
    case TLPC_C128:
      return run<Cx<double>>(dtype, x, A, zi, y, dx, dA, dzi, B, T, n, flags, strategy, chunk_len,
                             ws, ws_bytes, stream, backward, neg);
    default:
      return TLPC_EDTYPE;
  }
}

}  // namespace
}  // namespace tlpc

using namespace tlpc;

#define TLPC_EXPORT extern "C" __attribute__((visibility("default")))

TLPC_EXPORT const char* tlpc_error_string(int code) {
  switch (code) {
    case TLPC_OK: return "ok";
    case TLPC_EINVAL: return "invalid argument";
    case TLPC_EDTYPE: return "unsupported dtype";
    case TLPC_EWORKSPACE: return "workspace too small";
    case TLPC_ELAUNCH: return "kernel launch failed";
    case TLPC_EORDER: return "unsupported filter order";
    case TLPC_ESCRATCH: return "sizes exceed the static scratch of tlpc_lpc_forward_noalloc";
    case TLPC_EDEVICE: return "no current HIP device, or its scratch address is unavailable";
    default: return "unknown error";
  }
}

TLPC_EXPORT int tlpc_abi_version(void) { return 1; }

TLPC_EXPORT size_t tlpc_lpc_forward_workspace(int dtype, int64_t B, int64_t T, int32_t N,
                                              int32_t flags, int32_t strategy, int32_t chunk_len) {
  if (check_common(dtype, B, T, N) != TLPC_OK) return 0;
  return make_plan(dtype, B, T, N, flags, strategy, chunk_len, false).total;
}

TLPC_EXPORT int tlpc_lpc_forward(int dtype, const void* x, const void* A, const void* zi, void* y,
                                 int64_t B, int64_t T, int32_t N, int32_t flags, int32_t strategy,
                                 int32_t chunk_len, void* workspace, size_t workspace_bytes,
                                 void* hip_stream) {
  return run_dtype(dtype, x, A, zi, y, nullptr, nullptr, nullptr, B, T, N,
                   flags & TLPC_TIME_INVARIANT, strategy, chunk_len, workspace, workspace_bytes,
                   hip_stream, false);
}

TLPC_EXPORT int tlpc_lpc_forward_noalloc(int dtype, const void* x, const void* A, const void* zi,
                                         void* y, int64_t B, int64_t T, int32_t N, int32_t flags,
                                         int32_t strategy, int32_t chunk_len, void* hip_stream) {
  int rc = check_common(dtype, B, T, N);
  if (rc != TLPC_OK) return rc;
  const int32_t f = flags & TLPC_TIME_INVARIANT;
  const size_t need = make_plan(dtype, B, T, N, f, strategy, chunk_len, false).total;
  if (need > TLPC_NOALLOC_SCRATCH_BYTES) return TLPC_ESCRATCH;
  void* ws = nullptr;
  if (need > 0) {
    rc = scratch_address(&ws);
    if (rc != TLPC_OK) return rc;
  }
  return run_dtype(dtype, x, A, zi, y, nullptr, nullptr, nullptr, B, T, N, f, strategy, chunk_len,
                   ws, need > 0 ? TLPC_NOALLOC_SCRATCH_BYTES : 0, hip_stream, false);
}

TLPC_EXPORT size_t tlpc_lpc_backward_workspace(int dtype, int64_t B, int64_t T, int32_t N,
                                               int32_t flags, int32_t strategy, int32_t chunk_len) {
  if (check_common(dtype, B, T, N) != TLPC_OK) return 0;
  return make_plan(dtype, B, T, N, flags, strategy, chunk_len, true).total;
}

TLPC_EXPORT int tlpc_lpc_backward(int dtype, const void* g, const void* A, const void* zi,
                                  const void* y, void* dx, void* dA, void* dzi, int64_t B,
                                  int64_t T, int32_t N, int32_t flags, int32_t strategy,
                                  int32_t chunk_len, void* workspace, size_t workspace_bytes,
                                  void* hip_stream) {
  return run_dtype(dtype, g, A, zi, const_cast<void*>(y), dx, dA, dzi, B, T, N, flags, strategy,
                   chunk_len, workspace, workspace_bytes, hip_stream, true);
}

TLPC_EXPORT int tlpc_lpc_recompute_supported(int dtype, int64_t B, int64_t T, int32_t N,
                                             int32_t flags, int32_t strategy, int32_t chunk_len) {
  if (check_common(dtype, B, T, N) != TLPC_OK) return 0;
  return recompute_ok(dtype, B, T, N, flags, strategy, chunk_len) ? 1 : 0;
}

TLPC_EXPORT int tlpc_lpc_backward_recompute(int dtype, const void* g, const void* A, const void* zi,
                                            const void* x, const void* fwd_workspace,
                                            size_t fwd_workspace_bytes, void* dx, void* dA, void* dzi,
                                            int64_t B, int64_t T, int32_t N, int32_t flags,
                                            int32_t strategy, int32_t chunk_len, void* workspace,
                                            size_t workspace_bytes, void* hip_stream) {
  const int rc = check_common(dtype, B, T, N);
  if (rc != TLPC_OK) return rc;
  const Recompute r{x, fwd_workspace, fwd_workspace_bytes};
  if (dtype == TLPC_F32)
    return run<float>(dtype, g, A, zi, nullptr, dx, dA, dzi, B, T, N, flags & ~TLPC_REUSE_SUMMARY, strategy,
                      chunk_len, workspace, workspace_bytes, hip_stream, true, 0, &r);
  if (dtype == TLPC_F64)
    return run<double>(dtype, g, A, zi, nullptr, dx, dA, dzi, B, T, N, flags & ~TLPC_REUSE_SUMMARY, strategy,
                       chunk_len, workspace, workspace_bytes, hip_stream, true, 0, &r);
  return TLPC_EINVAL;
}

TLPC_EXPORT int tlpc_lpc_plan(int dtype, int64_t B, int64_t T, int32_t N, int32_t flags,
                              int32_t strategy, int32_t chunk_len, int32_t backward,
                              int32_t* strategy_out, int32_t* chunk_len_out) {
  const int rc = check_common(dtype, B, T, N);
  if (rc != TLPC_OK) return rc;
  if (!strategy_out || !chunk_len_out) return TLPC_EINVAL;
  const Plan p = make_plan(dtype, B, T, N, backward ? flags : (flags & TLPC_TIME_INVARIANT),
                           strategy, chunk_len, backward != 0);
  *strategy_out = p.strategy;
  *chunk_len_out = p.L;
  return TLPC_OK;
}

TLPC_EXPORT size_t tlpc_lpc_summary_bytes(int dtype, int64_t B, int64_t T, int32_t N,
                                          int32_t flags, int32_t strategy, int32_t chunk_len) {
  if (check_common(dtype, B, T, N) != TLPC_OK) return 0;
  const Plan f = make_plan(dtype, B, T, N, flags & TLPC_TIME_INVARIANT, strategy, chunk_len, false);
  const Plan b = make_plan(dtype, B, T, N, flags, strategy, chunk_len, true);
  if (!f.c16 || !b.c16 || f.L != b.L || f.summ_bytes != b.summ_bytes) return 0;
  return f.summ_bytes;
}

TLPC_EXPORT int tlpc_lpc_ragged_plan(int64_t K, const int64_t* x_off, const int64_t* a_off,
                                     const int64_t* z_off, const int64_t* lengths, int32_t N,
                                     int32_t dtype, int32_t chunk_len, int64_t* table,
                                     int64_t* table_len, int64_t* G_out, int64_t* max_chunks,
                                     int32_t* chunk_len_out) {
  if (K < 0 || N < 1 || !lengths || !table_len || !G_out || !max_chunks || !chunk_len_out)
    return TLPC_EINVAL;
  const int np = compiled_order(N);
  if (np == 0) return TLPC_EORDER;
  int64_t total = 0;
  for (int64_t k = 0; k < K; ++k) {
    if (lengths[k] < 1) return TLPC_EINVAL;
    total += lengths[k];
  }
  if (elem_size(dtype) == 0) return TLPC_EDTYPE;
  const int L = pick_chunk(K > 0 ? K : 1, K > 0 ? (total + K - 1) / K : 1, np, chunk_len,
                           multi_lane_rerun(dtype, N, np));
  int64_t G = 0, maxP = 0;
  for (int64_t k = 0; k < K; ++k) {
    const int64_t P = (lengths[k] + L - 1) / L;
    G += P;
    maxP = P > maxP ? P : maxP;
  }
  *G_out = G;
  *max_chunks = maxP;
  *chunk_len_out = L;
  *table_len = 5 * K + 1 + G;
  if (!table) return TLPC_OK;
  if (!x_off || !a_off || !z_off) return TLPC_EINVAL;
  int64_t c = 0;
  for (int64_t k = 0; k < K; ++k) {
    table[k] = x_off[k];
    table[K + k] = a_off[k];
    table[2 * K + k] = z_off[k];
    table[3 * K + k] = lengths[k];
    table[4 * K + k] = c;
    const int64_t P = (lengths[k] + L - 1) / L;
    for (int64_t j = 0; j < P; ++j) table[5 * K + 1 + c + j] = k;
    c += P;
  }
  table[5 * K] = c;
  return TLPC_OK;
}

TLPC_EXPORT size_t tlpc_lpc_ragged_workspace(int dtype, int64_t G, int32_t N) {
  const int np = compiled_order(N);
  if (wide_size(dtype) == 0 || np == 0 || G < 0) return 0;
  return ragged_workspace(wide_size(dtype), G, np);
}

TLPC_EXPORT int tlpc_lpc_ragged_forward(int dtype, const void* x, const void* A, const void* zi,
                                        void* y, const int64_t* table, int64_t K, int64_t G,
                                        int64_t max_chunks, int32_t N, int32_t chunk_len,
                                        void* workspace, size_t workspace_bytes,
                                        void* hip_stream) {
  return run_ragged_dtype(dtype, x, A, zi, y, nullptr, nullptr, nullptr, table, K, G, max_chunks,
                          N, 0, chunk_len, workspace, workspace_bytes, hip_stream, false);
}

TLPC_EXPORT int tlpc_lpc_ragged_backward(int dtype, const void* g, const void* A, const void* zi,
                                         const void* y, void* dx, void* dA, void* dzi,
                                         const int64_t* table, int64_t K, int64_t G,
                                         int64_t max_chunks, int32_t N, int32_t flags,
                                         int32_t chunk_len, void* workspace,
                                         size_t workspace_bytes, void* hip_stream) {
  return run_ragged_dtype(dtype, g, A, zi, const_cast<void*>(y), dx, dA, dzi, table, K, G,
                          max_chunks, N, flags, chunk_len, workspace, workspace_bytes, hip_stream,
                          true);
}

TLPC_EXPORT size_t tlpc_scan_workspace(int dtype, int64_t B, int64_t T, int32_t flags,
                                       int32_t chunk_len) {
  if (check_common(dtype, B, T, 1) != TLPC_OK) return 0;
  return make_plan(dtype, B, T, 1, flags, TLPC_STRATEGY_AUTO, chunk_len,
                   (flags & (TLPC_NEED_DX | TLPC_NEED_DA | TLPC_NEED_DZI)) != 0)
      .total;
}

TLPC_EXPORT int tlpc_scan(int dtype, const void* x, const void* w, const void* init, void* h,
                          int64_t B, int64_t T, int32_t chunk_len, void* workspace,
                          size_t workspace_bytes, void* hip_stream) {
  return run_dtype(dtype, x, w, init, h, nullptr, nullptr, nullptr, B, T, 1, 0,
                   TLPC_STRATEGY_AUTO, chunk_len, workspace, workspace_bytes, hip_stream, false, 1);
}

TLPC_EXPORT int tlpc_scan_backward(int dtype, const void* g, const void* w, const void* init,
                                   const void* h, void* dx, void* dw, void* dinit, int64_t B,
                                   int64_t T, int32_t flags, int32_t chunk_len, void* workspace,
                                   size_t workspace_bytes, void* hip_stream) {
  return run_dtype(dtype, g, w, init, const_cast<void*>(h), dx, dw, dinit, B, T, 1,
                   flags & ~TLPC_TIME_INVARIANT, TLPC_STRATEGY_AUTO, chunk_len, workspace,
                   workspace_bytes, hip_stream, true, 1);
}
