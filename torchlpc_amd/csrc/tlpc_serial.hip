// One thread per signal, no time parallelism. Used for orders above the largest
// compiled chunked order (32) and on request (TLPC_STRATEGY_SERIAL). It follows
// the reference loop (torchlpc/csrc/scan_cpu.cpp:112-120) with y[t-i] read back
// from the output, and the adjoint in direct form with shifted coefficients
// lambda[t] = g[t] - sum_i conj(A[t+i,i]) lambda[t+i] (torchlpc/core.py:194-213).
#include "tlpc_common.h"
#include "tlpc_kernels.h"

namespace tlpc {

template <typename S>
__global__ __launch_bounds__(64) void fwd_serial(ChunkArgs<S> a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.B) return;
  const S* xb = a.x + b * a.T;
  const S* Ab = a.A + b * a.a_bstride;
  const S* zib = a.zi ? a.zi + b * a.n : nullptr;
  S* yb = a.y + b * a.T;
  for (int64_t t = 0; t < a.T; ++t) {
    S acc = xb[t];
    const S* row = Ab + t * a.a_tstride;
    for (int i = 1; i <= a.n; ++i) {
      const int64_t u = t - i;
      const S yv = u >= 0 ? yb[u] : (zib ? zib[-u - 1] : Num<S>::zero());
      acc = Num<S>::fms(acc, coef(a, row[i - 1]), yv);
    }
    yb[t] = acc;
  }
}

// lambda is kept in `lam` ([B,T], dx or workspace).
template <typename S>
__global__ __launch_bounds__(64) void bwd_serial(ChunkArgs<S> a, S* lam) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.B) return;
  const S* gb = a.x + b * a.T;
  const S* Ab = a.A + b * a.a_bstride;
  const S* yb = a.y + b * a.T;
  const S* zib = a.zi ? a.zi + b * a.n : nullptr;
  S* lb = lam + b * a.T;
  for (int64_t t = a.T - 1; t >= 0; --t) {
    S acc = gb[t];
    for (int i = 1; i <= a.n && t + i < a.T; ++i)
      acc = Num<S>::fms(acc, Num<S>::conj(coef(a, Ab[(t + i) * a.a_tstride + i - 1])), lb[t + i]);
    lb[t] = acc;
  }
  if (a.dzi) {
    for (int k = 0; k < a.n; ++k) {
      // lambda at position -1-k receives contributions from t = i-1-k >= 0
      S acc = Num<S>::zero();
      for (int i = k + 1; i <= a.n; ++i) {
        const int64_t t = i - 1 - k;
        if (t < a.T) acc = Num<S>::fms(acc, Num<S>::conj(coef(a, Ab[t * a.a_tstride + i - 1])), lb[t]);
      }
      a.dzi[b * a.n + k] = acc;
    }
  }
  auto yval = [&](int64_t u) -> S {
    return u >= 0 ? Num<S>::conj(yb[u]) : (zib ? Num<S>::conj(zib[-u - 1]) : Num<S>::zero());
  };
  if (a.dA) {
    S* db = a.dA + b * a.T * a.n;
    for (int64_t t = 0; t < a.T; ++t)
      for (int i = 1; i <= a.n; ++i) db[t * a.n + i - 1] = coef(a, -(lb[t] * yval(t - i)));
  }
  if (a.da_ti) {
    for (int i = 1; i <= a.n; ++i) {
      S acc = Num<S>::zero();
      for (int64_t t = 0; t < a.T; ++t) acc = Num<S>::fms(acc, lb[t], yval(t - i));
      a.da_ti[b * a.n + i - 1] = coef(a, acc);
    }
  }
}

template <typename S>
__global__ __launch_bounds__(256) void fill_zero_kernel(S* p, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = Num<S>::zero();
}

template <typename S>
int launch_serial_fwd(ChunkArgs<S> a, hipStream_t st) {
  hipLaunchKernelGGL((fwd_serial<S>), dim3((unsigned)((a.B + 63) / 64)), dim3(64), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <typename S>
int launch_serial_bwd(ChunkArgs<S> a, hipStream_t st) {
  // lambda lives in dx, or in the workspace (`carry`, sized [B,T] for this strategy)
  S* lam = a.dx ? a.dx : reinterpret_cast<S*>(a.carry);
  hipLaunchKernelGGL((bwd_serial<S>), dim3((unsigned)((a.B + 63) / 64)), dim3(64), 0, st, a, lam);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <typename S>
int launch_fill_zero(S* p, int64_t count, hipStream_t st) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL((fill_zero_kernel<S>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

#define TLPC_INST_SERIAL(S)                                       \
  template int launch_serial_fwd<S>(ChunkArgs<S>, hipStream_t);   \
  template int launch_serial_bwd<S>(ChunkArgs<S>, hipStream_t);   \
  template int launch_fill_zero<S>(S*, int64_t, hipStream_t);

TLPC_INST_SERIAL(float)
TLPC_INST_SERIAL(double)
TLPC_INST_SERIAL(Cx<float>)
TLPC_INST_SERIAL(Cx<double>)

}  // namespace tlpc
