/*
 * CPU oracle for the all-pole (LPC) filter: TEST INFRASTRUCTURE ONLY.
 *
 * This file restates the reference CPU path so that the GPU kernels can be
 * checked against it. Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it, and only as a checker / timed CPU baseline. The
 * product (torchlpc_amd) never calls it.
 *
 * Followed reference code (DiffAPF/torchlpc 0.8):
 *   lpc forward   torchlpc/csrc/scan_cpu.cpp:80-122 (lpc_cpu_core) and :143-164 (lpc_cpu):
 *                 out = cat(zi.flip(1), x); for t: acc = out[t]; for i = 0..N-1:
 *                 acc -= a[t*N+i] * out[t-i-1]; out[t] = acc.  Same loop and
 *                 accumulation order, parallel over B only (scan_cpu.cpp:106).
 *   scan          torchlpc/csrc/scan_cpu.cpp:62-77: h = h * w + x.
 *   lpc backward  torchlpc/core.py:186-229: shifted coefficients
 *                 shifted[s, i-1] = A[s-N+i, i-1] (s in [0, T+N)), the forward loop on the
 *                 time-reversed, N-zero-padded gradient with conj(shifted), then
 *                 dzi = last N outputs, dx = the rest reversed,
 *                 dA[t,i-1] = conj(y[t-i]) * (-dx[t]).
 *   scan backward torchlpc/recurrence.py:92-126 (same construction with order 1).
 *
 * Compiled with plain gcc flags (no -ffast-math, no FMA contraction beyond what
 * the target allows by default), like the reference extension (setup.py:28-30).
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define DEFINE_LPC(NAME, T_)                                                               \
  /* y[B,T] = lpc(x[B,T], A[B,T,N], zi[B,N]); scan_cpu.cpp:80-122,143-164 */              \
  void NAME(const T_* x, const T_* A, const T_* zi, T_* y, int64_t B, int64_t T,         \
            int64_t N) {                                                                   \
    _Pragma("omp parallel for schedule(static, 1)") for (int64_t b = 0; b < B; ++b) {      \
      T_* out = (T_*)malloc(sizeof(T_) * (size_t)(T + N));                                 \
      for (int64_t k = 0; k < N; ++k) out[k] = zi[b * N + (N - 1 - k)];                    \
      memcpy(out + N, x + b * T, sizeof(T_) * (size_t)T);                                  \
      T_* o = out + N;                                                                     \
      const T_* a = A + b * T * N;                                                         \
      for (int64_t t = 0; t < T; ++t) {                                                    \
        T_ acc = o[t];                                                                     \
        for (int64_t i = 0; i < N; ++i) acc -= a[t * N + i] * o[t - i - 1];                \
        o[t] = acc;                                                                        \
      }                                                                                    \
      memcpy(y + b * T, o, sizeof(T_) * (size_t)T);                                        \
      free(out);                                                                           \
    }                                                                                      \
  }

DEFINE_LPC(oracle_lpc_f32, float)
DEFINE_LPC(oracle_lpc_f64, double)
DEFINE_LPC(oracle_lpc_c64, float complex)
DEFINE_LPC(oracle_lpc_c128, double complex)

#define DEFINE_SCAN(NAME, T_)                                                              \
  /* h[B,T]: h = h * w + x, h(-1) = init[b]; scan_cpu.cpp:62-77 */                         \
  void NAME(const T_* x, const T_* w, const T_* init, T_* h, int64_t B, int64_t T) {       \
    _Pragma("omp parallel for schedule(static, 1)") for (int64_t b = 0; b < B; ++b) {      \
      T_ s = init[b];                                                                      \
      for (int64_t t = 0; t < T; ++t) {                                                    \
        s = s * w[b * T + t] + x[b * T + t];                                               \
        h[b * T + t] = s;                                                                  \
      }                                                                                    \
    }                                                                                      \
  }

DEFINE_SCAN(oracle_scan_f32, float)
DEFINE_SCAN(oracle_scan_f64, double)
DEFINE_SCAN(oracle_scan_c64, float complex)
DEFINE_SCAN(oracle_scan_c128, double complex)

static inline float conjf_(float v) { return v; }
static inline double conj_(double v) { return v; }
static inline float complex conjc64_(float complex v) { return conjf(v); }
static inline double complex conjc128_(double complex v) { return conj(v); }

/*
 * Backward composition of torchlpc/core.py:186-229, for gradient g[B,T] and the
 * saved (A, zi, y). Writes dx[B,T], dA[B,T,N], dzi[B,N] (each may be NULL).
 * The filter pass is the lpc loop above on the reversed sequence of length T+N
 * (zi needs a gradient) with coefficients conj(shifted A) reversed in time.
 */
#define DEFINE_LPC_BWD(NAME, LPC, T_, CONJ)                                                \
  void NAME(const T_* g, const T_* A, const T_* zi, const T_* y, T_* dx, T_* dA, T_* dzi,  \
            int64_t B, int64_t T, int64_t N) {                                             \
    const int64_t S = T + N;                                                               \
    T_* rg = (T_*)calloc((size_t)(B * S), sizeof(T_));                                     \
    T_* rA = (T_*)calloc((size_t)(B * S * N), sizeof(T_));                                 \
    T_* z0 = (T_*)calloc((size_t)(B * N), sizeof(T_));                                     \
    T_* out = (T_*)malloc(sizeof(T_) * (size_t)(B * S));                                   \
    _Pragma("omp parallel for") for (int64_t b = 0; b < B; ++b) {                          \
      /* padded gradient [0]*N ++ g, flipped: rg[u] = g[T-1-u] for u < T */               \
      for (int64_t u = 0; u < T; ++u) rg[b * S + u] = g[b * T + (T - 1 - u)];              \
      /* shifted[s, i-1] = A[s-N+i, i-1]; reversed in time and conjugated */               \
      for (int64_t s = 0; s < S; ++s)                                                      \
        for (int64_t i = 1; i <= N; ++i) {                                                 \
          const int64_t t = s - N + i;                                                     \
          if (t >= 0 && t < T)                                                             \
            rA[(b * S + (S - 1 - s)) * N + (i - 1)] = CONJ(A[(b * T + t) * N + (i - 1)]);  \
        }                                                                                  \
    }                                                                                      \
    LPC(rg, rA, z0, out, B, S, N);                                                         \
    _Pragma("omp parallel for") for (int64_t b = 0; b < B; ++b) {                          \
      const T_* o = out + b * S;                                                           \
      if (dzi)                                                                             \
        for (int64_t k = 0; k < N; ++k) dzi[b * N + k] = o[T + k];                         \
      if (dx)                                                                              \
        for (int64_t t = 0; t < T; ++t) dx[b * T + t] = o[T - 1 - t];                      \
      if (dA)                                                                              \
        for (int64_t t = 0; t < T; ++t)                                                    \
          for (int64_t i = 1; i <= N; ++i) {                                               \
            const int64_t u = t - i;                                                       \
            const T_ yv = u >= 0 ? y[b * T + u] : zi[b * N + (-u - 1)];                    \
            dA[(b * T + t) * N + (i - 1)] = CONJ(yv) * (-o[T - 1 - t]);                    \
          }                                                                                \
    }                                                                                      \
    free(rg);                                                                              \
    free(rA);                                                                              \
    free(z0);                                                                              \
    free(out);                                                                             \
  }

DEFINE_LPC_BWD(oracle_lpc_backward_f32, oracle_lpc_f32, float, conjf_)
DEFINE_LPC_BWD(oracle_lpc_backward_f64, oracle_lpc_f64, double, conj_)
DEFINE_LPC_BWD(oracle_lpc_backward_c64, oracle_lpc_c64, float complex, conjc64_)
DEFINE_LPC_BWD(oracle_lpc_backward_c128, oracle_lpc_c128, double complex, conjc128_)

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
