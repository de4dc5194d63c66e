// PyTorch operator layer over the C ABI (include/torchlpc_amd.h).
//
// Registers the namespace `torchlpc_amd` with the ops the Python autograd
// Functions call. The GPU implementations validate with TORCH_CHECK, allocate the
// outputs and the workspace from the caching allocator, and launch on the current
// HIP stream; no call synchronises, so the ops can be captured in a hipGraph.
// CPU tensors are rejected with an error: the filter runs on the GPU only.
//
// Reference op boundary: TORCH_LIBRARY(torchlpc) schemas at
// torchlpc/csrc/scan_cpu.cpp:166-170 (lpc(x, A, zi), scan(input, weights, initials)).
#include <ATen/ATen.h>
#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <algorithm>
#include <tuple>

#include "../../include/torchlpc_amd.h"

namespace {

int dtype_code(const at::Tensor& t) {
  switch (t.scalar_type()) {
    case at::kFloat: return TLPC_F32;
    case at::kDouble: return TLPC_F64;
    case at::kComplexFloat: return TLPC_C64;
    case at::kComplexDouble: return TLPC_C128;
    default:
      TORCH_CHECK(false, "torchlpc_amd: unsupported dtype ", t.scalar_type(),
                  " (float32, float64, complex64 and complex128 are supported)");
  }
  return -1;
}

void check_rc(int rc, const char* what) {
  TORCH_CHECK(rc == TLPC_OK, "torchlpc_amd: ", what, " failed: ", tlpc_error_string(rc), " (",
              rc, ")");
}

void* stream_of(const at::Tensor& t) {
  return reinterpret_cast<void*>(c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

at::Tensor workspace(const at::Tensor& like, size_t bytes) {
  return at::empty({(int64_t)bytes}, like.options().dtype(at::kByte));
}

// Coefficient layout resolution. Returns the coefficient tensor to hand to the
// kernels and the time-invariance flag: a 2-D a[B,N] and a 3-D A whose time
// stride is 0 (a time broadcast, e.g. a[:, None, :].expand(-1, T, -1)) are read
// as one row per signal without materialising [B,T,N].
std::pair<at::Tensor, bool> coef_view(const at::Tensor& A, int64_t T) {
  if (A.dim() == 2) return {A.contiguous(), true};
  if (A.size(1) > 1 && A.stride(1) == 0) return {A.select(1, 0).contiguous(), true};
  return {A.contiguous(), false};
}

void check_inputs(const at::Tensor& x, const at::Tensor& A, const at::Tensor& zi, const char* op) {
  TORCH_CHECK(x.is_cuda(), "torchlpc_amd::", op,
              ": tensors must be on a HIP (MI355X) device; CPU execution is not provided");
  TORCH_CHECK(x.dim() == 2, "torchlpc_amd::", op, ": x must be 2-D [B, T], got ", x.sizes());
  TORCH_CHECK(A.dim() == 3 || A.dim() == 2, "torchlpc_amd::", op,
              ": A must be [B, T, N] or [B, N], got ", A.sizes());
  TORCH_CHECK(A.scalar_type() == x.scalar_type() && zi.scalar_type() == x.scalar_type(),
              "torchlpc_amd::", op, ": x, A and zi must share a dtype");
  TORCH_CHECK(A.device() == x.device() && zi.device() == x.device(), "torchlpc_amd::", op,
              ": x, A and zi must be on the same device");
  TORCH_CHECK(A.size(0) == x.size(0), "torchlpc_amd::", op, ": batch of A and x differ");
  if (A.dim() == 3)
    TORCH_CHECK(A.size(1) == x.size(1), "torchlpc_amd::", op, ": time of A and x differ");
  const int64_t N = A.size(A.dim() - 1);
  TORCH_CHECK(zi.dim() == 2 && zi.size(0) == x.size(0) && zi.size(1) == N, "torchlpc_amd::", op,
              ": zi must be [B, N] = [", x.size(0), ", ", N, "], got ", zi.sizes());
  TORCH_CHECK(N >= 1, "torchlpc_amd::", op, ": filter order must be >= 1");
}

// Forward; with `keep` the workspace is sized for the backward as well and
// returned when the backward can reuse its chunk summaries (else an empty tensor).
std::tuple<at::Tensor, at::Tensor> lpc_forward_impl(const at::Tensor& x_, const at::Tensor& A_,
                                                    const at::Tensor& zi_, int64_t strategy,
                                                    int64_t chunk_len, bool keep) {
  check_inputs(x_, A_, zi_, "lpc");
  c10::DeviceGuard guard(x_.device());
  const auto x = x_.contiguous();
  const auto zi = zi_.contiguous();
  const int64_t B = x.size(0), T = x.size(1), N = A_.size(A_.dim() - 1);
  auto [A, ti] = coef_view(A_, T);
  auto y = at::empty_like(x);
  const int dt = dtype_code(x);
  const int flags = ti ? TLPC_TIME_INVARIANT : 0;
  size_t wsb = tlpc_lpc_forward_workspace(dt, B, T, (int32_t)N, flags, (int32_t)strategy,
                                          (int32_t)chunk_len);
  keep = keep && tlpc_lpc_summary_bytes(dt, B, T, (int32_t)N, flags, (int32_t)strategy,
                                        (int32_t)chunk_len) > 0;
  if (keep) {
    const int all = TLPC_NEED_DX | TLPC_NEED_DA | TLPC_NEED_DZI | flags;
    wsb = std::max(wsb, tlpc_lpc_backward_workspace(dt, B, T, (int32_t)N, all, (int32_t)strategy,
                                                    (int32_t)chunk_len));
  }
  auto ws = workspace(x, wsb);
  check_rc(tlpc_lpc_forward(dt, x.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(), B, T,
                            (int32_t)N, flags, (int32_t)strategy, (int32_t)chunk_len,
                            wsb ? ws.data_ptr() : nullptr, wsb, stream_of(x)),
           "lpc forward");
  return {y, keep ? ws : at::empty({0}, x.options().dtype(at::kByte))};
}

at::Tensor lpc_forward_hip(const at::Tensor& x, const at::Tensor& A, const at::Tensor& zi,
                           int64_t strategy, int64_t chunk_len) {
  return std::get<0>(lpc_forward_impl(x, A, zi, strategy, chunk_len, false));
}

std::tuple<at::Tensor, at::Tensor> lpc_keep_hip(const at::Tensor& x, const at::Tensor& A,
                                                const at::Tensor& zi, int64_t strategy,
                                                int64_t chunk_len) {
  return lpc_forward_impl(x, A, zi, strategy, chunk_len, true);
}

std::tuple<at::Tensor, at::Tensor, at::Tensor> lpc_backward_hip(
    const at::Tensor& g_, const at::Tensor& A_, const at::Tensor& zi_, const at::Tensor& y_,
    bool need_x, bool need_a, bool need_zi, int64_t strategy, int64_t chunk_len,
    const c10::optional<at::Tensor>& summary) {
  check_inputs(g_, A_, zi_, "lpc_backward");
  TORCH_CHECK(y_.sizes() == g_.sizes(), "torchlpc_amd::lpc_backward: y and grad shapes differ");
  c10::DeviceGuard guard(g_.device());
  const auto g = g_.contiguous();
  const auto zi = zi_.contiguous();
  const auto y = y_.contiguous();
  const int64_t B = g.size(0), T = g.size(1), N = A_.size(A_.dim() - 1);
  auto [A, ti] = coef_view(A_, T);
  const bool full_da = ti && A_.dim() == 3;  // gradient of a time-broadcast A is per sample
  at::Tensor dx = need_x ? at::empty_like(g) : at::empty({0}, g.options());
  at::Tensor dA = need_a ? (A_.dim() == 3 ? at::empty({B, T, N}, g.options())
                                          : at::empty({B, N}, g.options()))
                         : at::empty({0}, g.options());
  at::Tensor dzi = need_zi ? at::empty({B, N}, g.options()) : at::empty({0}, g.options());
  int flags = (need_x ? TLPC_NEED_DX : 0) | (need_a ? TLPC_NEED_DA : 0) |
              (need_zi ? TLPC_NEED_DZI : 0) | (ti ? TLPC_TIME_INVARIANT : 0) |
              (full_da ? TLPC_DA_FULL : 0);
  if (!(need_x || need_a || need_zi)) return {dx, dA, dzi};
  const int dt = dtype_code(g);
  const size_t wsb = tlpc_lpc_backward_workspace(dt, B, T, (int32_t)N, flags, (int32_t)strategy,
                                                 (int32_t)chunk_len);
  at::Tensor ws;
  if (summary.has_value() && summary->numel() > 0) {
    // the forward's workspace (lpc_keep): its leading summaries are reused
    TORCH_CHECK(summary->device() == g.device() && summary->scalar_type() == at::kByte &&
                    summary->is_contiguous() && (size_t)summary->numel() >= wsb &&
                    tlpc_lpc_summary_bytes(dt, B, T, (int32_t)N, flags, (int32_t)strategy,
                                           (int32_t)chunk_len) > 0,
                "torchlpc_amd::lpc_backward: `summary` does not come from lpc_keep with the same "
                "shapes, strategy and chunk_len");
    ws = *summary;
    flags |= TLPC_REUSE_SUMMARY;
  } else {
    ws = workspace(g, wsb);
  }
  check_rc(tlpc_lpc_backward(dt, g.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(),
                             need_x ? dx.data_ptr() : nullptr, need_a ? dA.data_ptr() : nullptr,
                             need_zi ? dzi.data_ptr() : nullptr, B, T, (int32_t)N, flags,
                             (int32_t)strategy, (int32_t)chunk_len, wsb ? ws.data_ptr() : nullptr,
                             wsb ? (size_t)ws.numel() : 0, stream_of(g)),
           "lpc backward");
  return {dx, dA, dzi};
}

// Forward that returns its whole workspace: the chunk carries an adjoint without
// a saved y rebuilds y from (lpc_backward_recompute).
std::tuple<at::Tensor, at::Tensor> lpc_carries_hip(const at::Tensor& x_, const at::Tensor& A_,
                                                   const at::Tensor& zi_, int64_t strategy,
                                                   int64_t chunk_len) {
  check_inputs(x_, A_, zi_, "lpc_carries");
  TORCH_CHECK(A_.dim() == 3, "torchlpc_amd::lpc_carries: A must be time-varying [B, T, N]");
  c10::DeviceGuard guard(x_.device());
  const auto x = x_.contiguous();
  const auto zi = zi_.contiguous();
  const auto A = A_.contiguous();
  const int64_t B = x.size(0), T = x.size(1), N = A.size(2);
  const int dt = dtype_code(x);
  TORCH_CHECK(tlpc_lpc_recompute_supported(dt, B, T, (int32_t)N, TLPC_NEED_DX | TLPC_NEED_DA | TLPC_NEED_DZI,
                                           (int32_t)strategy, (int32_t)chunk_len) == 1,
              "torchlpc_amd::lpc_carries: the backward cannot rebuild y for this shape, dtype, "
              "strategy and chunk length (real data, N <= 4, chunked scan, chunks <= 256)");
  auto y = at::empty_like(x);
  const size_t wsb = tlpc_lpc_forward_workspace(dt, B, T, (int32_t)N, 0, (int32_t)strategy, (int32_t)chunk_len);
  auto ws = workspace(x, wsb);
  check_rc(tlpc_lpc_forward(dt, x.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(), B, T, (int32_t)N, 0,
                            (int32_t)strategy, (int32_t)chunk_len, wsb ? ws.data_ptr() : nullptr, wsb,
                            stream_of(x)),
           "lpc forward");
  return {y, ws};
}

std::tuple<at::Tensor, at::Tensor, at::Tensor> lpc_backward_recompute_hip(
    const at::Tensor& g_, const at::Tensor& A_, const at::Tensor& zi_, const at::Tensor& x_,
    const at::Tensor& fws, bool need_x, bool need_a, bool need_zi, int64_t strategy,
    int64_t chunk_len) {
  check_inputs(g_, A_, zi_, "lpc_backward_recompute");
  TORCH_CHECK(A_.dim() == 3, "torchlpc_amd::lpc_backward_recompute: A must be time-varying [B, T, N]");
  TORCH_CHECK(x_.sizes() == g_.sizes() && x_.scalar_type() == g_.scalar_type() && x_.device() == g_.device(),
              "torchlpc_amd::lpc_backward_recompute: x must be the forward input, shaped like grad");
  c10::DeviceGuard guard(g_.device());
  const auto g = g_.contiguous();
  const auto zi = zi_.contiguous();
  const auto x = x_.contiguous();
  const auto A = A_.contiguous();
  const int64_t B = g.size(0), T = g.size(1), N = A.size(2);
  at::Tensor dx = need_x ? at::empty_like(g) : at::empty({0}, g.options());
  at::Tensor dA = need_a ? at::empty({B, T, N}, g.options()) : at::empty({0}, g.options());
  at::Tensor dzi = need_zi ? at::empty({B, N}, g.options()) : at::empty({0}, g.options());
  if (!(need_x || need_a || need_zi)) return {dx, dA, dzi};
  const int flags = (need_x ? TLPC_NEED_DX : 0) | (need_a ? TLPC_NEED_DA : 0) | (need_zi ? TLPC_NEED_DZI : 0);
  const int dt = dtype_code(g);
  const size_t fwb = tlpc_lpc_forward_workspace(dt, B, T, (int32_t)N, 0, (int32_t)strategy, (int32_t)chunk_len);
  TORCH_CHECK(fws.device() == g.device() && fws.scalar_type() == at::kByte && fws.is_contiguous() &&
                  (size_t)fws.numel() >= fwb,
              "torchlpc_amd::lpc_backward_recompute: the workspace does not come from lpc_carries "
              "with the same shapes, strategy and chunk_len");
  const size_t wsb = tlpc_lpc_backward_workspace(dt, B, T, (int32_t)N, flags, (int32_t)strategy,
                                                 (int32_t)chunk_len);
  auto ws = workspace(g, wsb);
  check_rc(tlpc_lpc_backward_recompute(dt, g.data_ptr(), A.data_ptr(), zi.data_ptr(), x.data_ptr(),
                                       fws.data_ptr(), (size_t)fws.numel(), need_x ? dx.data_ptr() : nullptr,
                                       need_a ? dA.data_ptr() : nullptr, need_zi ? dzi.data_ptr() : nullptr, B,
                                       T, (int32_t)N, flags, (int32_t)strategy, (int32_t)chunk_len,
                                       wsb ? ws.data_ptr() : nullptr, wsb, stream_of(g)),
           "lpc backward (rebuilding y)");
  return {dx, dA, dzi};
}

void check_scan(const at::Tensor& x, const at::Tensor& w, const at::Tensor& init, const char* op) {
  TORCH_CHECK(x.is_cuda(), "torchlpc_amd::", op,
              ": tensors must be on a HIP (MI355X) device; CPU execution is not provided");
  TORCH_CHECK(x.dim() == 2, "torchlpc_amd::", op, ": input must be 2-D [B, T]");
  TORCH_CHECK(w.sizes() == x.sizes(), "torchlpc_amd::", op, ": weights must match the input shape");
  TORCH_CHECK(init.dim() == 1 && init.size(0) == x.size(0), "torchlpc_amd::", op,
              ": initials must be [B]");
  TORCH_CHECK(w.scalar_type() == x.scalar_type() && init.scalar_type() == x.scalar_type(),
              "torchlpc_amd::", op, ": input, weights and initials must share a dtype");
  TORCH_CHECK(w.device() == x.device() && init.device() == x.device(), "torchlpc_amd::", op,
              ": tensors must be on the same device");
}

at::Tensor scan_hip(const at::Tensor& x_, const at::Tensor& w_, const at::Tensor& init_,
                    int64_t chunk_len) {
  check_scan(x_, w_, init_, "scan");
  c10::DeviceGuard guard(x_.device());
  const auto x = x_.contiguous(), w = w_.contiguous(), init = init_.contiguous();
  const int64_t B = x.size(0), T = x.size(1);
  auto h = at::empty_like(x);
  const int dt = dtype_code(x);
  const size_t wsb = tlpc_scan_workspace(dt, B, T, 0, (int32_t)chunk_len);
  auto ws = workspace(x, wsb);
  check_rc(tlpc_scan(dt, x.data_ptr(), w.data_ptr(), init.data_ptr(), h.data_ptr(), B, T,
                     (int32_t)chunk_len, wsb ? ws.data_ptr() : nullptr, wsb, stream_of(x)),
           "scan");
  return h;
}

std::tuple<at::Tensor, at::Tensor, at::Tensor> scan_backward_hip(
    const at::Tensor& g_, const at::Tensor& w_, const at::Tensor& init_, const at::Tensor& h_,
    bool need_x, bool need_w, bool need_init, int64_t chunk_len) {
  check_scan(g_, w_, init_, "scan_backward");
  c10::DeviceGuard guard(g_.device());
  const auto g = g_.contiguous(), w = w_.contiguous(), init = init_.contiguous(),
             h = h_.contiguous();
  const int64_t B = g.size(0), T = g.size(1);
  at::Tensor dx = need_x ? at::empty_like(g) : at::empty({0}, g.options());
  at::Tensor dw = need_w ? at::empty_like(g) : at::empty({0}, g.options());
  at::Tensor dinit = need_init ? at::empty({B}, g.options()) : at::empty({0}, g.options());
  if (!(need_x || need_w || need_init)) return {dx, dw, dinit};
  const int flags = (need_x ? TLPC_NEED_DX : 0) | (need_w ? TLPC_NEED_DA : 0) |
                    (need_init ? TLPC_NEED_DZI : 0);
  const int dt = dtype_code(g);
  const size_t wsb = tlpc_scan_workspace(dt, B, T, flags, (int32_t)chunk_len);
  auto ws = workspace(g, wsb);
  check_rc(tlpc_scan_backward(dt, g.data_ptr(), w.data_ptr(), init.data_ptr(), h.data_ptr(),
                              need_x ? dx.data_ptr() : nullptr, need_w ? dw.data_ptr() : nullptr,
                              need_init ? dinit.data_ptr() : nullptr, B, T, flags,
                              (int32_t)chunk_len, wsb ? ws.data_ptr() : nullptr, wsb,
                              stream_of(g)),
           "scan backward");
  return {dx, dw, dinit};
}

// ---------------------------------------------------------------------------
// Ragged batches of one order (offsets and lengths in a device table built by
// the host plan below).
// Returns the device-ready table and (G, max chunks, chunk length, x extent,
// A extent, zi extent): the extents are the end of the furthest signal in the
// packed x / y, A / dA and zi / dzi tensors, which the launches check.
std::tuple<at::Tensor, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t> lpc_ragged_plan_cpu(
    const at::Tensor& x_off, const at::Tensor& a_off, const at::Tensor& z_off,
    const at::Tensor& lengths, int64_t N, at::ScalarType dtype, int64_t chunk_len) {
  for (const auto* t : {&x_off, &a_off, &z_off, &lengths})
    TORCH_CHECK(t->device().is_cpu() && t->scalar_type() == at::kLong && t->dim() == 1 &&
                    t->numel() == lengths.numel(),
                "torchlpc_amd::lpc_ragged_plan: offsets and lengths must be 1-D int64 CPU "
                "tensors of one length");
  const auto xo = x_off.contiguous(), ao = a_off.contiguous(), zo = z_off.contiguous(),
             ln = lengths.contiguous();
  const int64_t K = ln.numel();
  const int dt = dtype_code(at::empty({0}, at::TensorOptions().dtype(dtype)));
  int64_t tlen = 0, G = 0, maxP = 0;
  int32_t L = 0;
  check_rc(tlpc_lpc_ragged_plan(K, nullptr, nullptr, nullptr, ln.data_ptr<int64_t>(), (int32_t)N,
                                dt, (int32_t)chunk_len, nullptr, &tlen, &G, &maxP, &L),
           "ragged plan");
  auto table = at::empty({tlen}, at::TensorOptions().dtype(at::kLong));
  check_rc(tlpc_lpc_ragged_plan(K, xo.data_ptr<int64_t>(), ao.data_ptr<int64_t>(),
                                zo.data_ptr<int64_t>(), ln.data_ptr<int64_t>(), (int32_t)N, dt,
                                (int32_t)chunk_len, table.data_ptr<int64_t>(), &tlen, &G, &maxP,
                                &L),
           "ragged plan");
  int64_t xe = 0, ae = 0, ze = 0;
  const int64_t* xp = xo.data_ptr<int64_t>();
  const int64_t* ap = ao.data_ptr<int64_t>();
  const int64_t* zp = zo.data_ptr<int64_t>();
  const int64_t* lp = ln.data_ptr<int64_t>();
  for (int64_t k = 0; k < K; ++k) {
    TORCH_CHECK(xp[k] >= 0 && ap[k] >= 0 && zp[k] >= 0,
                "torchlpc_amd::lpc_ragged_plan: offsets must be non-negative");
    xe = std::max(xe, xp[k] + lp[k]);
    ae = std::max(ae, ap[k] + lp[k] * N);
    ze = std::max(ze, zp[k] + N);
  }
  return {table, G, maxP, (int64_t)L, xe, ae, ze};
}

void check_ragged(const at::Tensor& x, const at::Tensor& A, const at::Tensor& zi,
                  const at::Tensor& table, int64_t K, int64_t G, at::IntArrayRef extents,
                  const char* op) {
  TORCH_CHECK(x.is_cuda(), "torchlpc_amd::", op,
              ": tensors must be on a HIP (MI355X) device; CPU execution is not provided");
  TORCH_CHECK(x.dim() == 1 && A.dim() == 1 && zi.dim() == 1, "torchlpc_amd::", op,
              ": packed x, A and zi must be 1-D");
  TORCH_CHECK(x.is_contiguous() && A.is_contiguous() && zi.is_contiguous(), "torchlpc_amd::", op,
              ": packed tensors must be contiguous");
  TORCH_CHECK(A.scalar_type() == x.scalar_type() && zi.scalar_type() == x.scalar_type(),
              "torchlpc_amd::", op, ": x, A and zi must share a dtype");
  TORCH_CHECK(table.device() == x.device() && A.device() == x.device() && zi.device() == x.device(),
              "torchlpc_amd::", op, ": tensors must be on one device");
  TORCH_CHECK(table.scalar_type() == at::kLong && table.is_contiguous(), "torchlpc_amd::", op,
              ": the plan table must be a contiguous int64 device tensor");
  TORCH_CHECK(table.numel() == 5 * K + 1 + G, "torchlpc_amd::", op, ": table size mismatch");
  TORCH_CHECK(extents.size() == 3, "torchlpc_amd::", op,
              ": extents must be (x, A, zi) as returned by lpc_ragged_plan");
  TORCH_CHECK(x.numel() >= extents[0] && A.numel() >= extents[1] && zi.numel() >= extents[2],
              "torchlpc_amd::", op, ": packed tensors (x ", x.numel(), ", A ", A.numel(), ", zi ",
              zi.numel(), " elements) are smaller than the plan's signals need (", extents[0], ", ",
              extents[1], ", ", extents[2], ")");
}

// Ragged forward into y; returns the workspace, which holds the chunk summaries
// a later lpc_ragged_backward_ may be handed (`kept`).
at::Tensor ragged_forward(const at::Tensor& x, const at::Tensor& A, const at::Tensor& zi,
                          const at::Tensor& table, at::Tensor& y, int64_t K, int64_t G, int64_t maxP,
                          int64_t N, int64_t chunk_len, at::IntArrayRef extents, const char* op) {
  check_ragged(x, A, zi, table, K, G, extents, op);
  TORCH_CHECK(y.sizes() == x.sizes() && y.scalar_type() == x.scalar_type() && y.is_contiguous(),
              "torchlpc_amd::", op, ": y must be a contiguous tensor like x");
  c10::DeviceGuard guard(x.device());
  const int dt = dtype_code(x);
  const size_t wsb = tlpc_lpc_ragged_workspace(dt, G, (int32_t)N);
  auto ws = workspace(x, wsb);
  check_rc(tlpc_lpc_ragged_forward(dt, x.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(),
                                   table.data_ptr<int64_t>(), K, G, maxP, (int32_t)N,
                                   (int32_t)chunk_len, wsb ? ws.data_ptr() : nullptr, wsb,
                                   stream_of(x)),
           "ragged forward");
  return ws;
}

void lpc_ragged_hip(const at::Tensor& x, const at::Tensor& A, const at::Tensor& zi,
                    const at::Tensor& table, at::Tensor& y, int64_t K, int64_t G, int64_t maxP,
                    int64_t N, int64_t chunk_len, at::IntArrayRef extents) {
  ragged_forward(x, A, zi, table, y, K, G, maxP, N, chunk_len, extents, "lpc_ragged_");
}

at::Tensor lpc_ragged_keep_hip(const at::Tensor& x, const at::Tensor& A, const at::Tensor& zi,
                               const at::Tensor& table, at::Tensor& y, int64_t K, int64_t G,
                               int64_t maxP, int64_t N, int64_t chunk_len, at::IntArrayRef extents) {
  return ragged_forward(x, A, zi, table, y, K, G, maxP, N, chunk_len, extents, "lpc_ragged_keep_");
}

void lpc_ragged_backward_hip(const at::Tensor& g, const at::Tensor& A, const at::Tensor& zi,
                             const at::Tensor& y, const at::Tensor& table, at::Tensor& dx,
                             at::Tensor& dA, at::Tensor& dzi, int64_t K, int64_t G, int64_t maxP,
                             int64_t N, int64_t chunk_len, bool need_x, bool need_a,
                             bool need_zi, at::IntArrayRef extents,
                             const c10::optional<at::Tensor>& kept) {
  check_ragged(g, A, zi, table, K, G, extents, "lpc_ragged_backward_");
  TORCH_CHECK(y.sizes() == g.sizes() && y.is_contiguous(),
              "torchlpc_amd::lpc_ragged_backward_: y must be a contiguous tensor like g");
  if (need_x) TORCH_CHECK(dx.sizes() == g.sizes() && dx.is_contiguous(), "lpc_ragged_backward_: dx");
  if (need_a) TORCH_CHECK(dA.sizes() == A.sizes() && dA.is_contiguous(), "lpc_ragged_backward_: dA");
  if (need_zi) TORCH_CHECK(dzi.sizes() == zi.sizes() && dzi.is_contiguous(), "lpc_ragged_backward_: dzi");
  if (!(need_x || need_a || need_zi)) return;
  c10::DeviceGuard guard(g.device());
  const int dt = dtype_code(g);
  int flags = (need_x ? TLPC_NEED_DX : 0) | (need_a ? TLPC_NEED_DA : 0) |
              (need_zi ? TLPC_NEED_DZI : 0);
  const size_t wsb = tlpc_lpc_ragged_workspace(dt, G, (int32_t)N);
  at::Tensor ws;
  if (kept.has_value() && kept->defined()) {
    // the workspace of lpc_ragged_keep_ for this table: its summaries are reused
    TORCH_CHECK(kept->scalar_type() == at::kByte && kept->is_contiguous() && kept->device() == g.device() &&
                    (size_t)kept->numel() >= wsb,
                "torchlpc_amd::lpc_ragged_backward_: `kept` must be the workspace lpc_ragged_keep_ returned "
                "for this plan group (",
                wsb, " bytes)");
    ws = *kept;
    flags |= TLPC_REUSE_SUMMARY;
  } else {
    ws = workspace(g, wsb);
  }
  check_rc(tlpc_lpc_ragged_backward(dt, g.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(),
                                    need_x ? dx.data_ptr() : nullptr,
                                    need_a ? dA.data_ptr() : nullptr,
                                    need_zi ? dzi.data_ptr() : nullptr,
                                    table.data_ptr<int64_t>(), K, G, maxP, (int32_t)N, flags,
                                    (int32_t)chunk_len, wsb ? ws.data_ptr() : nullptr, wsb,
                                    stream_of(g)),
           "ragged backward");
}

[[noreturn]] void cpu_unsupported(const char* op) {
  TORCH_CHECK(false, "torchlpc_amd::", op,
              ": CPU tensors are not supported; move x, A and zi to the MI355X device (.cuda())");
  throw;  // unreachable
}

at::Tensor lpc_cpu(const at::Tensor&, const at::Tensor&, const at::Tensor&, int64_t, int64_t) {
  cpu_unsupported("lpc");
}
std::tuple<at::Tensor, at::Tensor> lpc_carries_cpu(const at::Tensor&, const at::Tensor&, const at::Tensor&,
                                                   int64_t, int64_t) {
  cpu_unsupported("lpc_carries");
}
std::tuple<at::Tensor, at::Tensor, at::Tensor> lpc_backward_recompute_cpu(const at::Tensor&, const at::Tensor&,
                                                                          const at::Tensor&, const at::Tensor&,
                                                                          const at::Tensor&, bool, bool, bool,
                                                                          int64_t, int64_t) {
  cpu_unsupported("lpc_backward_recompute");
}
std::tuple<at::Tensor, at::Tensor> lpc_keep_cpu(const at::Tensor&, const at::Tensor&,
                                                const at::Tensor&, int64_t, int64_t) {
  cpu_unsupported("lpc_keep");
}
std::tuple<at::Tensor, at::Tensor, at::Tensor> lpc_backward_cpu(
    const at::Tensor&, const at::Tensor&, const at::Tensor&, const at::Tensor&, bool, bool, bool,
    int64_t, int64_t, const c10::optional<at::Tensor>&) {
  cpu_unsupported("lpc_backward");
}
void lpc_ragged_cpu(const at::Tensor&, const at::Tensor&, const at::Tensor&, const at::Tensor&,
                    at::Tensor&, int64_t, int64_t, int64_t, int64_t, int64_t, at::IntArrayRef) {
  cpu_unsupported("lpc_ragged_");
}
at::Tensor lpc_ragged_keep_cpu(const at::Tensor&, const at::Tensor&, const at::Tensor&, const at::Tensor&,
                               at::Tensor&, int64_t, int64_t, int64_t, int64_t, int64_t, at::IntArrayRef) {
  cpu_unsupported("lpc_ragged_keep_");
}
void lpc_ragged_backward_cpu(const at::Tensor&, const at::Tensor&, const at::Tensor&,
                             const at::Tensor&, const at::Tensor&, at::Tensor&, at::Tensor&,
                             at::Tensor&, int64_t, int64_t, int64_t, int64_t, int64_t, bool, bool,
                             bool, at::IntArrayRef, const c10::optional<at::Tensor>&) {
  cpu_unsupported("lpc_ragged_backward_");
}
at::Tensor scan_cpu(const at::Tensor&, const at::Tensor&, const at::Tensor&, int64_t) {
  cpu_unsupported("scan");
}
std::tuple<at::Tensor, at::Tensor, at::Tensor> scan_backward_cpu(const at::Tensor&,
                                                                 const at::Tensor&,
                                                                 const at::Tensor&,
                                                                 const at::Tensor&, bool, bool,
                                                                 bool, int64_t) {
  cpu_unsupported("scan_backward");
}

}  // namespace

TORCH_LIBRARY(torchlpc_amd, m) {
  m.def("lpc(Tensor x, Tensor A, Tensor zi, int strategy=0, int chunk_len=0) -> Tensor");
  m.def(
      "lpc_keep(Tensor x, Tensor A, Tensor zi, int strategy=0, int chunk_len=0) -> "
      "(Tensor, Tensor)");
  m.def(
      "lpc_backward(Tensor grad_y, Tensor A, Tensor zi, Tensor y, bool need_x, bool need_a, "
      "bool need_zi, int strategy=0, int chunk_len=0, Tensor? summary=None) -> "
      "(Tensor, Tensor, Tensor)");
  m.def(
      "lpc_carries(Tensor x, Tensor A, Tensor zi, int strategy=0, int chunk_len=0) -> "
      "(Tensor, Tensor)");
  m.def(
      "lpc_backward_recompute(Tensor grad_y, Tensor A, Tensor zi, Tensor x, Tensor fwd_workspace, "
      "bool need_x, bool need_a, bool need_zi, int strategy=0, int chunk_len=0) -> "
      "(Tensor, Tensor, Tensor)");
  m.def(
      "lpc_ragged_plan(Tensor x_off, Tensor a_off, Tensor z_off, Tensor lengths, int N, "
      "ScalarType dtype, int chunk_len=0) -> (Tensor, int, int, int, int, int, int)");
  m.def(
      "lpc_ragged_(Tensor x, Tensor A, Tensor zi, Tensor table, Tensor(a!) y, int K, int G, "
      "int max_chunks, int N, int chunk_len, int[] extents) -> ()");
  m.def(
      "lpc_ragged_keep_(Tensor x, Tensor A, Tensor zi, Tensor table, Tensor(a!) y, int K, int G, "
      "int max_chunks, int N, int chunk_len, int[] extents) -> Tensor");
  m.def(
      "lpc_ragged_backward_(Tensor g, Tensor A, Tensor zi, Tensor y, Tensor table, Tensor(a!) dx, "
      "Tensor(b!) dA, Tensor(c!) dzi, int K, int G, int max_chunks, int N, int chunk_len, "
      "bool need_x, bool need_a, bool need_zi, int[] extents, Tensor(d!)? kept=None) -> ()");
  m.def("scan(Tensor x, Tensor w, Tensor init, int chunk_len=0) -> Tensor");
  m.def(
      "scan_backward(Tensor grad, Tensor w, Tensor init, Tensor h, bool need_x, bool need_w, "
      "bool need_init, int chunk_len=0) -> (Tensor, Tensor, Tensor)");
}

TORCH_LIBRARY_IMPL(torchlpc_amd, CUDA, m) {
  m.impl("lpc", &lpc_forward_hip);
  m.impl("lpc_keep", &lpc_keep_hip);
  m.impl("lpc_backward", &lpc_backward_hip);
  m.impl("lpc_carries", &lpc_carries_hip);
  m.impl("lpc_backward_recompute", &lpc_backward_recompute_hip);
  m.impl("lpc_ragged_", &lpc_ragged_hip);
  m.impl("lpc_ragged_keep_", &lpc_ragged_keep_hip);
  m.impl("lpc_ragged_backward_", &lpc_ragged_backward_hip);
  m.impl("scan", &scan_hip);
  m.impl("scan_backward", &scan_backward_hip);
}

TORCH_LIBRARY_IMPL(torchlpc_amd, CPU, m) {
  m.impl("lpc", &lpc_cpu);
  m.impl("lpc_keep", &lpc_keep_cpu);
  m.impl("lpc_backward", &lpc_backward_cpu);
  m.impl("lpc_carries", &lpc_carries_cpu);
  m.impl("lpc_backward_recompute", &lpc_backward_recompute_cpu);
  m.impl("lpc_ragged_plan", &lpc_ragged_plan_cpu);
  m.impl("lpc_ragged_", &lpc_ragged_cpu);
  m.impl("lpc_ragged_keep_", &lpc_ragged_keep_cpu);
  m.impl("lpc_ragged_backward_", &lpc_ragged_backward_cpu);
  m.impl("scan", &scan_cpu);
  m.impl("scan_backward", &scan_backward_cpu);
}
