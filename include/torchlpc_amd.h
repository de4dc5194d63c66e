/*
 * torchlpc_amd C ABI: time-varying all-pole (LPC) filter on MI355X (gfx950).
 *
 *   y[b,t] = x[b,t] - sum_{i=1..N} A[b,t,i-1] * y[b,t-i],   y[b,-k] = zi[b,k-1]
 *
 * Every entry point is a plain C function over device pointers and sizes.
 * The caller owns all buffers (inputs, outputs and the workspace); the library
 * never allocates device memory, never synchronises and launches only on the
 * stream it is given, so every call can be recorded into a hipGraph.
 *
 * Layouts (all row-major, contiguous, element type selected by `dtype`):
 *   x, y, g, dx : [B, T]
 *   A, dA       : [B, T, N]      (time-varying coefficients)
 *   a, da       : [B, N]         (time-invariant coefficients, TLPC_TIME_INVARIANT)
 *   zi, dzi     : [B, N]         zi[b,k] = y[b,-1-k]  (same orientation as the reference)
 * Ragged (packed) batches describe signal k by offsets instead of B/T:
 *   x/y/g/dx at x + off_t[k], length len[k];  A/dA at A + off_a[k], [len[k], N_k];
 *   zi/dzi at zi + off_z[k], N_k values.
 *
 * Return value: TLPC_OK (0) or a negative TLPC_E* code; tlpc_error_string()
 * names it. A failed kernel launch is reported as TLPC_ELAUNCH.
 */
#ifndef TORCHLPC_AMD_H
#define TORCHLPC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types */
#define TLPC_F32 0
#define TLPC_F64 1
#define TLPC_C64 2  /* complex float, interleaved (re, im) */
#define TLPC_C128 3 /* complex double, interleaved (re, im) */

/* status codes */
#define TLPC_OK 0
#define TLPC_EINVAL -1     /* bad size, null pointer, bad flag */
#define TLPC_EDTYPE -2     /* unsupported element type */
#define TLPC_EWORKSPACE -3 /* workspace smaller than the matching *_workspace() query */
#define TLPC_ELAUNCH -4    /* hipGetLastError() reported a launch failure */
#define TLPC_EORDER -5     /* unsupported filter order */
#define TLPC_ESCRATCH -6   /* tlpc_lpc_forward_noalloc: sizes need more than its static scratch */
#define TLPC_EDEVICE -7    /* no current HIP device, or its scratch address could not be read */

/* flags for the backward entry points: which gradients to write */
#define TLPC_NEED_DX 1
#define TLPC_NEED_DA 2
#define TLPC_NEED_DZI 4
/* flag for forward/backward: A is time-invariant, given as a[B, N] */
#define TLPC_TIME_INVARIANT 8
/* with TLPC_TIME_INVARIANT in the backward: write the per-sample dA[B,T,N]
   (the gradient of a time-broadcast A) instead of the reduced da[B,N] */
#define TLPC_DA_FULL 16
/* backward: the first tlpc_lpc_summary_bytes() bytes of the workspace still hold
   what tlpc_lpc_forward wrote there for the same dtype, B, T, N, TIME_INVARIANT
   flag, strategy and chunk_len (the chunk transition matrices depend on A only);
   the backward reads them instead of recomputing them. Ignored when
   tlpc_lpc_summary_bytes() is 0. */
#define TLPC_REUSE_SUMMARY 32

/* strategy for tlpc_lpc_forward / tlpc_lpc_backward (`strategy` argument) */
#define TLPC_STRATEGY_AUTO 0    /* library heuristic */
#define TLPC_STRATEGY_CHUNKED 1 /* multi-kernel chunked scan (any B) */
#define TLPC_STRATEGY_SIGNAL 2  /* one workgroup per signal walking LDS tiles (float32, N = 16) */
#define TLPC_STRATEGY_SERIAL 3  /* one thread per signal, no time parallelism */

const char* tlpc_error_string(int code);
int tlpc_abi_version(void);

/*
 * Forward filter. Replaces the reference op torchlpc::lpc(x, A, zi)
 * (torchlpc/csrc/scan_cpu.cpp:143-164 CPU, torchlpc/csrc/cuda/lpc.cu:115-175 CUDA,
 * schema at scan_cpu.cpp:169). `chunk_len` 0 lets the library choose.
 * B == 0 is a no-op. With T == 0 the per-sample buffers (x, a per-sample A, y;
 * in the backward g, y, dx and a per-sample dA) hold nothing and may be NULL.
 */
size_t tlpc_lpc_forward_workspace(int dtype, int64_t B, int64_t T, int32_t N, int32_t flags,
                                  int32_t strategy, int32_t chunk_len);
int tlpc_lpc_forward(int dtype, const void* x, const void* A, const void* zi, void* y, int64_t B,
                     int64_t T, int32_t N, int32_t flags, int32_t strategy, int32_t chunk_len,
                     void* workspace, size_t workspace_bytes, void* hip_stream);

/*
 * Allocation-free forward: tlpc_lpc_forward with the workspace taken from a
 * static device array of TLPC_NOALLOC_SCRATCH_BYTES per GPU instead of the
 * caller. No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It calls no hipMalloc / hipFree and no
 * allocating library routine. Sizes whose tlpc_lpc_forward_workspace() exceeds
 * TLPC_NOALLOC_SCRATCH_BYTES are rejected with TLPC_ESCRATCH before anything is
 * launched. The array is held by every GPU the library runs on until the
 * process ends, and every call on one GPU shares it: calls that overlap on one
 * GPU (concurrent streams, or host threads) race and must be serialised by the
 * caller. Nothing in it needs to start at zero (every kernel writes the scratch
 * it reads within the same call). The first call on a device, with that device
 * current, reads the array's device address (hipGetSymbolAddress) and keeps it.
 */
#define TLPC_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tlpc_lpc_forward_noalloc(int dtype, const void* x, const void* A, const void* zi, void* y,
                             int64_t B, int64_t T, int32_t N, int32_t flags, int32_t strategy,
                             int32_t chunk_len, void* hip_stream);

/*
 * Fused adjoint. Replaces the backward composition of LPC.backward
 * (torchlpc/core.py:186-229): one reverse-time pass gives
 *   dx[t]      = lambda[t] = g[t] - sum_i conj(A[t+i,i]) lambda[t+i]
 *   dA[t,i-1]  = -lambda[t] * conj(y[t-i])          (y[-k] = zi[k-1])
 *   dzi[k]     = lambda[-1-k]
 * With TLPC_TIME_INVARIANT, `dA` receives da[B,N] = -sum_t lambda[t] conj(y[t-i]),
 * or the per-sample dA[B,T,N] when TLPC_DA_FULL is also set.
 * Gradients whose TLPC_NEED_* flag is clear may be passed as NULL.
 */
size_t tlpc_lpc_backward_workspace(int dtype, int64_t B, int64_t T, int32_t N, int32_t flags,
                                   int32_t strategy, int32_t chunk_len);
int tlpc_lpc_backward(int dtype, const void* g, const void* A, const void* zi, const void* y,
                      void* dx, void* dA, void* dzi, int64_t B, int64_t T, int32_t N,
                      int32_t flags, int32_t strategy, int32_t chunk_len, void* workspace,
                      size_t workspace_bytes, void* hip_stream);

/*
 * The backward of tlpc_lpc_backward without the forward's output y: the adjoint
 * rerun rebuilds y chunk by chunk from the forward input x and the chunk
 * carries that tlpc_lpc_forward left in its workspace (`fwd_workspace`, the
 * whole forward workspace of the same dtype, B, T, N, strategy and chunk_len,
 * untouched since). The results equal tlpc_lpc_backward's given the forward's
 * y. `workspace` is the backward's own (tlpc_lpc_backward_workspace). For long
 * signals this keeps B*P*N float64 carries between the passes instead of B*T
 * samples of y. Served when tlpc_lpc_recompute_supported() returns 1: real
 * dtypes, time-varying A, N <= 4, the chunked strategy with chunks of at most
 * 256 samples (TLPC_EINVAL otherwise). Replaces the saved output of
 * torchlpc/core.py:179-184 and its read in core.py:222-227.
 */
int tlpc_lpc_recompute_supported(int dtype, int64_t B, int64_t T, int32_t N, int32_t flags,
                                 int32_t strategy, int32_t chunk_len);
int tlpc_lpc_backward_recompute(int dtype, const void* g, const void* A, const void* zi,
                                const void* x, const void* fwd_workspace, size_t fwd_workspace_bytes,
                                void* dx, void* dA, void* dzi, int64_t B, int64_t T, int32_t N,
                                int32_t flags, int32_t strategy, int32_t chunk_len, void* workspace,
                                size_t workspace_bytes, void* hip_stream);

/*
 * The plan that tlpc_lpc_forward (backward = 0) or tlpc_lpc_backward (backward =
 * 1) runs for these arguments: the resolved strategy (TLPC_STRATEGY_*, never
 * AUTO) and the chunk length (0 for the per-signal strategy, whose tiles are
 * fixed). Host only, launches nothing. No reference counterpart (the reference
 * has one CUDA kernel per op, torchlpc/csrc/cuda/lpc.cu:115-175).
 */
int tlpc_lpc_plan(int dtype, int64_t B, int64_t T, int32_t N, int32_t flags, int32_t strategy,
                  int32_t chunk_len, int32_t backward, int32_t* strategy_out,
                  int32_t* chunk_len_out);

/*
 * Size of the leading workspace region that a forward and a backward with these
 * arguments share (see TLPC_REUSE_SUMMARY), 0 when the backward cannot reuse the
 * forward's summaries. A caller that wants the reuse gives the forward a
 * workspace of at least the backward's size and passes the same one to the
 * backward. No reference counterpart: the reference's backward
 * (torchlpc/core.py:186-229) recomputes everything from A.
 */
size_t tlpc_lpc_summary_bytes(int dtype, int64_t B, int64_t T, int32_t N, int32_t flags,
                              int32_t strategy, int32_t chunk_len);

/*
 * Ragged batches: K signals of one order N (N <= 32) with their own lengths,
 * packed back to back. No reference counterpart: the reference takes only
 * rectangular [B, T] batches (torchlpc/__init__.py:63-104); these entry points
 * give the same per-signal results as tlpc_lpc_forward / tlpc_lpc_backward on
 * each signal alone. Signal k reads x[x_off[k] .. + lengths[k]) (also the
 * offset of y, g, dx), A[a_off[k] .. + lengths[k]*N) as [lengths[k], N] rows
 * (also dA), zi[z_off[k] .. + N) (also dzi); offsets are in elements.
 *
 * tlpc_lpc_ragged_plan (host only) chooses the chunk length (chunk_len 0: the
 * library's choice) and fills `table` (int64, table_len entries; pass NULL to
 * query table_len, G and the chunk length first). The caller copies the table
 * to device memory and passes it, with K, G, max_chunks and the chunk length,
 * to the launches. Time-varying coefficients only.
 *
 * tlpc_lpc_ragged_backward takes TLPC_REUSE_SUMMARY in `flags` when its workspace
 * still holds what tlpc_lpc_ragged_forward wrote there for the same table, A,
 * N, dtype and chunk length: for float32 orders 8, 16 and 32 the backward then
 * reads the forward's chunk transition matrices (transposed) instead of forming
 * its own, and overwrites only the rest of the workspace, so the flag may be
 * given again to a later backward with the same workspace. It is ignored for
 * other orders and dtypes (the backward then forms its own summaries there).
 */
int tlpc_lpc_ragged_plan(int64_t K, const int64_t* x_off, const int64_t* a_off,
                         const int64_t* z_off, const int64_t* lengths, int32_t N, int32_t dtype,
                         int32_t chunk_len, int64_t* table, int64_t* table_len, int64_t* G,
                         int64_t* max_chunks, int32_t* chunk_len_out);
size_t tlpc_lpc_ragged_workspace(int dtype, int64_t G, int32_t N);
int tlpc_lpc_ragged_forward(int dtype, const void* x, const void* A, const void* zi, void* y,
                            const int64_t* table, int64_t K, int64_t G, int64_t max_chunks,
                            int32_t N, int32_t chunk_len, void* workspace, size_t workspace_bytes,
                            void* hip_stream);
int tlpc_lpc_ragged_backward(int dtype, const void* g, const void* A, const void* zi,
                             const void* y, void* dx, void* dA, void* dzi, const int64_t* table,
                             int64_t K, int64_t G, int64_t max_chunks, int32_t N, int32_t flags,
                             int32_t chunk_len, void* workspace, size_t workspace_bytes,
                             void* hip_stream);

/*
 * First-order recurrence h[t] = w[t] * h[t-1] + x[t], h[-1] = init[b] (the order-1
 * filter with A = -w). Replaces the reference op torchlpc::scan(input, weights,
 * initials) (torchlpc/csrc/scan_cpu.cpp:28-78,124-141, schema :168;
 * torchlpc/csrc/cuda/linear_recurrence.cu:232-289). `init` may be NULL (zero).
 */
size_t tlpc_scan_workspace(int dtype, int64_t B, int64_t T, int32_t flags, int32_t chunk_len);
int tlpc_scan(int dtype, const void* x, const void* w, const void* init, void* h, int64_t B,
              int64_t T, int32_t chunk_len, void* workspace, size_t workspace_bytes,
              void* hip_stream);
/*
 * Its adjoint (reference: Recurrence.backward, torchlpc/recurrence.py:92-126):
 *   mu[t] = g[t] + conj(w[t+1]) mu[t+1];  dx = mu;  dw[t] = mu[t] conj(h[t-1]);  dinit = conj(w[0]) mu[0]
 * `flags` selects outputs with TLPC_NEED_DX / TLPC_NEED_DA (dw) / TLPC_NEED_DZI (dinit);
 * query the workspace with the same flags.
 */
int tlpc_scan_backward(int dtype, const void* g, const void* w, const void* init, const void* h,
                       void* dx, void* dw, void* dinit, int64_t B, int64_t T, int32_t flags,
                       int32_t chunk_len, void* workspace, size_t workspace_bytes,
                       void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* TORCHLPC_AMD_H */
