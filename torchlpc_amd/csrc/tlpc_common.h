// Shared device-side helpers for the torchlpc_amd HIP kernels (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tlpc {

// Complex number with interleaved storage (matches torch complex64/complex128).
template <typename R>
struct alignas(2 * sizeof(R)) Cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ Cx<R> operator+(Cx<R> a, Cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ __forceinline__ Cx<R> operator-(Cx<R> a, Cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__device__ __forceinline__ Cx<R> operator-(Cx<R> a) { return {-a.re, -a.im}; }
template <typename R>
__device__ __forceinline__ Cx<R> operator*(Cx<R> a, Cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Scalar traits: zero, one, conj, and the multiply-subtract acc - a*b.
template <typename S>
struct Num {
  static __device__ __forceinline__ S zero() { return S(0); }
  static __device__ __forceinline__ S one() { return S(1); }
  static __device__ __forceinline__ S conj(S a) { return a; }
  static __device__ __forceinline__ S fms(S acc, S a, S b) { return acc - a * b; }
  static __device__ __forceinline__ S fma(S acc, S a, S b) { return acc + a * b; }
};
template <typename R>
struct Num<Cx<R>> {
  using S = Cx<R>;
  static __device__ __forceinline__ S zero() { return {R(0), R(0)}; }
  static __device__ __forceinline__ S one() { return {R(1), R(0)}; }
  static __device__ __forceinline__ S conj(S a) { return {a.re, -a.im}; }
  static __device__ __forceinline__ S fms(S acc, S a, S b) {
    return {acc.re - a.re * b.re + a.im * b.im, acc.im - a.re * b.im - a.im * b.re};
  }
  static __device__ __forceinline__ S fma(S acc, S a, S b) {
    return {acc.re + a.re * b.re - a.im * b.im, acc.im + a.re * b.im + a.im * b.re};
  }
};

// Wide type used for chunk summaries and carries: the basis responses of a chunk
// accumulate rounding over the whole chunk, so single-precision data keeps them
// in double precision; double-precision data stays as it is.
template <typename S>
struct Wide {
  using type = S;
};
template <>
struct Wide<float> {
  using type = double;
};
template <>
struct Wide<Cx<float>> {
  using type = Cx<double>;
};

template <typename D, typename S>
__device__ __forceinline__ D convert(S v) {
  return (D)v;
}
template <>
__device__ __forceinline__ Cx<double> convert<Cx<double>, Cx<float>>(Cx<float> v) {
  return {(double)v.re, (double)v.im};
}
template <>
__device__ __forceinline__ Cx<float> convert<Cx<float>, Cx<double>>(Cx<double> v) {
  return {(float)v.re, (float)v.im};
}
template <>
__device__ __forceinline__ Cx<float> convert<Cx<float>, Cx<float>>(Cx<float> v) {
  return v;
}
template <>
__device__ __forceinline__ Cx<double> convert<Cx<double>, Cx<double>>(Cx<double> v) {
  return v;
}

// Lane shuffle of a scalar (real or complex) within a wavefront.
__device__ __forceinline__ float shfl(float v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ double shfl(double v, int src) { return __shfl(v, src, 64); }
template <typename R>
__device__ __forceinline__ Cx<R> shfl(Cx<R> v, int src) {
  return {shfl(v.re, src), shfl(v.im, src)};
}

}  // namespace tlpc
