"""Host-side logic of the drop-in API (CPU only, no kernel runs).

Shapes and routing are checked on the meta device (the ops have fake kernels);
CPU tensors must raise a clear error rather than fall back to a CPU path.
"""
from __future__ import annotations

import pytest
import torch

import torchlpc_amd
from torchlpc_amd import LPC, linear_recurrence, sample_wise_lpc


def meta(*shape, dtype=torch.float32):
    return torch.empty(*shape, device="meta", dtype=dtype)


def test_extension_loaded_and_single_hip_runtime():
    assert torchlpc_amd.EXTENSION_LOADED, torchlpc_amd.LOAD_ERROR
    libs = torchlpc_amd.hip_runtimes_mapped()
    assert len(libs) == 1, libs


@pytest.mark.parametrize("B,T,N", [(2, 50, 3), (4, 7, 16), (1, 3, 5)])
def test_shapes_and_zf_convention(B, T, N):
    x, A, zi = meta(B, T), meta(B, T, N), meta(B, N)
    y = sample_wise_lpc(x, A, zi)
    assert y.shape == (B, T) and y.device.type == "meta"
    y, zf = sample_wise_lpc(x, A, zi, return_zf=True)
    # zf = y[:, -N:].flip(1): (B, N), or (B, T) when T < N (reference quirk)
    assert zf.shape == (B, min(N, T))


def test_time_invariant_shapes():
    y = sample_wise_lpc(meta(3, 40), meta(3, 4))
    assert y.shape == (3, 40)
    y = sample_wise_lpc(meta(3, 40), meta(3, 1, 4).expand(3, 40, 4))
    assert y.shape == (3, 40)


def test_order_one_routes_to_recurrence():
    y = sample_wise_lpc(meta(2, 30), meta(2, 30, 1), meta(2, 1))
    assert y.shape == (2, 30)
    assert y.grad_fn is None  # meta inputs without grad
    x = meta(2, 30).requires_grad_()
    y = sample_wise_lpc(x, meta(2, 30, 1))
    assert type(y.grad_fn).__name__ == "RecurrenceBackward"
    y = sample_wise_lpc(x, meta(2, 30, 3))
    assert type(y.grad_fn).__name__ == "LPCBackward"


def test_shape_asserts_match_reference():
    with pytest.raises(AssertionError):
        sample_wise_lpc(meta(2, 30), meta(3, 30, 2))
    with pytest.raises(AssertionError):
        sample_wise_lpc(meta(2, 30), meta(2, 31, 2))
    with pytest.raises(AssertionError):
        sample_wise_lpc(meta(2, 30), meta(2, 30, 2), meta(2, 3))
    with pytest.raises(AssertionError):
        linear_recurrence(meta(2, 30), meta(2, 31))


def test_cpu_tensors_raise_no_fallback():
    x, A, zi = torch.randn(2, 16), torch.randn(2, 16, 2) * 0.1, torch.zeros(2, 2)
    with pytest.raises(RuntimeError, match="CPU tensors are not supported"):
        sample_wise_lpc(x, A, zi)
    with pytest.raises(RuntimeError, match="CPU tensors are not supported"):
        linear_recurrence(torch.rand(2, 16), x)


def test_meta_backward_shapes():
    x = meta(3, 20).requires_grad_()
    A = meta(3, 20, 4).requires_grad_()
    zi = meta(3, 4).requires_grad_()
    y = LPC.apply(x, A, zi)
    gx, gA, gzi = torch.autograd.grad(y.sum(), (x, A, zi))
    assert gx.shape == x.shape and gA.shape == A.shape and gzi.shape == zi.shape
    a2 = meta(3, 4).requires_grad_()
    y = LPC.apply(x, a2, zi)
    (ga,) = torch.autograd.grad(y.sum(), (a2,))
    assert ga.shape == (3, 4)


def test_shifted_coefficients_identity():
    # S[s, i-1] = A[s-N+i, i-1]
    from torchlpc_amd.core import lagged_outputs, shifted_coefficients

    B, T, N = 2, 7, 3
    A = torch.arange(B * T * N, dtype=torch.float64).reshape(B, T, N) + 1
    S = shifted_coefficients(A)
    for s in range(T + N):
        for i in range(1, N + 1):
            t = s - N + i
            want = A[:, t, i - 1] if 0 <= t < T else torch.zeros(B, dtype=A.dtype)
            assert torch.equal(S[:, s, i - 1], want)
    y = torch.randn(B, T, dtype=torch.float64)
    zi = torch.randn(B, N, dtype=torch.float64)
    Y = lagged_outputs(y, zi, N)
    for t in range(T):
        for i in range(1, N + 1):
            want = y[:, t - i] if t - i >= 0 else zi[:, i - t - 1]
            assert torch.equal(Y[:, t, i - 1], want)


def test_autotune_table_roundtrip(tmp_path, monkeypatch):
    from torchlpc_amd import autotune

    monkeypatch.setenv("TORCHLPC_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.clear()
    x, A = meta(4, 1000), meta(4, 1000, 16)
    assert autotune.lookup("lpc_forward", x, A) == (0, 0)
    autotune.store("lpc_forward", x, A, (1, 256))
    assert autotune.lookup("lpc_forward", x, A) == (1, 256)
    assert (tmp_path / "tune.json").exists()
    # keyed on the exact length: a choice for T=1000 says nothing about T=1001
    assert autotune.lookup("lpc_forward", meta(4, 1001), meta(4, 1001, 16)) == (0, 0)
    autotune.clear()


def test_autotune_candidates():
    """Candidates: the heuristic, the per-signal kernels for float32 order 16, and
    chunked lengths that are multiples of the staging block (16, or 32 above
    order 16), within the signal."""
    from torchlpc_amd import autotune

    c16 = autotune.default_candidates(256, 48000, 16, torch.float32)
    assert c16[0] == (0, 0) and (2, 0) in c16
    assert all(L % 16 == 0 and 0 < L <= 48000 for s, L in c16 if s == 1) and len(c16) >= 5
    c32 = autotune.default_candidates(8, 5000, 32, torch.float64)
    assert (2, 0) not in c32 and all(L % 32 == 0 for s, L in c32 if s == 1)
    assert autotune.default_candidates(3, 0, 4, torch.float32) == [(0, 0)]
    assert (2, 0) in autotune.default_candidates(1024, 16384, 32, torch.float32, ti=True)
    assert (2, 0) not in autotune.default_candidates(1024, 16384, 32, torch.float32)


def test_meta_order16_keeps_forward_summaries():
    """A float32 order-16 forward that needs grad returns its workspace (fake kernel
    sized by the C ABI queries) and the backward accepts it."""
    from torchlpc_amd import _ext

    x = meta(2, 5000).requires_grad_()
    A = meta(2, 5000, 16).requires_grad_()
    zi = meta(2, 16)
    y, ws = torch.ops.torchlpc_amd.lpc_keep(x, A, zi)
    assert y.shape == (2, 5000) and ws.dtype == torch.uint8
    assert ws.numel() == _ext.keep_workspace_bytes(x, A, 0, 0) > 0
    dx, dA, dzi = torch.ops.torchlpc_amd.lpc_backward(y, A, zi, y, True, True, False, 0, 0, ws)
    assert dx.shape == x.shape and dA.shape == A.shape and dzi.numel() == 0
    y = LPC.apply(x, A, zi)
    gx, gA = torch.autograd.grad(y.sum(), (x, A))
    assert gx.shape == x.shape and gA.shape == A.shape
    # nothing to keep for float64 or a time-invariant a[B, N]
    assert _ext.keep_workspace_bytes(meta(2, 5000, dtype=torch.float64), meta(2, 5000, 16, dtype=torch.float64), 0, 0) == 0
    assert _ext.keep_workspace_bytes(x, meta(2, 16), 0, 0) == 0


def test_plan_query():
    """tlpc_lpc_plan: the per-signal kernels for a full float32 order-16 batch, the
    chunked scan elsewhere; time-invariant chunks of about sqrt(1.5 T) samples on
    long signals (a 1024-link chain, not 256 chunks of 4096 serial steps)."""
    from torchlpc_amd._ext import plan_of

    assert plan_of(meta(256, 48000), meta(256, 48000, 16)) == (2, 0)
    s, L = plan_of(meta(256, 48000, dtype=torch.float64), meta(256, 48000, 16, dtype=torch.float64))
    assert s == 1 and L > 0
    for N in (2, 32):
        s, L = plan_of(meta(4, 1 << 20), meta(4, N))
        assert s == 1 and 1254 <= L <= 1280 + 32, (N, L)
    # time-invariant order 16..32 over a full batch: the per-signal MFMA kernels
    assert plan_of(meta(1024, 16384), meta(1024, 32)) == (2, 0)
    assert plan_of(meta(1024, 16384), meta(1024, 12)) == (1, 352)
    assert plan_of(meta(4, 3000), meta(4, 3000, 16), strategy=1, chunk_len=100)[1] % 16 == 0


def test_recompute_mode_queries_and_selection(monkeypatch):
    """The y-rebuilding backward is served for long real time-varying signals of
    order up to 4 (chunks of at most 256 samples), and LPC.apply takes it for
    signals of RECOMPUTE_MIN_T samples and more: it saves x and the forward's
    workspace, not y, and the backward calls lpc_backward_recompute."""
    from torchlpc_amd import core
    from torchlpc_amd._ext import recompute_supported

    f64 = torch.float64
    assert recompute_supported(meta(4, 1 << 20), meta(4, 1 << 20, 2))
    assert recompute_supported(meta(4, 1 << 20, dtype=f64), meta(4, 1 << 20, 2, dtype=f64))
    assert not recompute_supported(meta(4, 5000), meta(4, 5000, 16))  # order 16
    assert not recompute_supported(meta(4, 5000), meta(4, 2))  # time-invariant
    assert not recompute_supported(meta(64, 1 << 20), meta(64, 1 << 20, 2))  # 2048-sample chunks
    assert recompute_supported(meta(64, 1 << 20), meta(64, 1 << 20, 2), 1, 256)
    assert not recompute_supported(meta(4, 5000, dtype=torch.complex64), meta(4, 5000, 2, dtype=torch.complex64))

    real = core.ops()
    calls = []

    class Spy:
        def __getattr__(self, name):
            return getattr(real, name)

        def lpc_carries(self, *a):
            calls.append("carries")
            return real.lpc_carries(*a)

        def lpc_backward_recompute(self, *a):
            calls.append("rebuild")
            return real.lpc_backward_recompute(*a)

    monkeypatch.setattr(core, "ops", lambda: Spy())
    T = core.RECOMPUTE_MIN_T
    x, A, zi = meta(2, T).requires_grad_(), meta(2, T, 2).requires_grad_(), meta(2, 2).requires_grad_()
    saved = []
    with torch.autograd.graph.saved_tensors_hooks(lambda t: saved.append(t) or t, lambda t: t):
        y = LPC.apply(x, A, zi)
    assert calls == ["carries"]
    assert any(t is x for t in saved) and not any(t is y for t in saved)
    gx, gA, gz = torch.autograd.grad(y.sum(), (x, A, zi))
    assert calls[-1] == "rebuild" and gA.shape == A.shape and gz.shape == zi.shape
    calls.clear()
    LPC.apply(meta(2, T - 16).requires_grad_(), meta(2, T - 16, 2), meta(2, 2))  # shorter: y is saved
    assert calls == []


def test_keep_decision_follows_the_callers_grad_mode(monkeypatch):
    """LPC.apply keeps the forward's workspace only when a backward can follow: the
    caller's grad mode on and an input that requires grad. The workspace is then a
    saved tensor (saved-tensor hooks see it) and reaches lpc_backward."""
    from torchlpc_amd import core

    real = core.ops()
    calls = []

    class Spy:
        def __getattr__(self, name):
            return getattr(real, name)

        def lpc(self, *a):
            calls.append("lpc")
            return real.lpc(*a)

        def lpc_keep(self, *a):
            calls.append("keep")
            return real.lpc_keep(*a)

        def lpc_backward(self, *a):
            calls.append(("backward", a[-1] is not None and a[-1].dtype == torch.uint8))
            return real.lpc_backward(*a)

    monkeypatch.setattr(core, "ops", lambda: Spy())
    x, A, zi = meta(2, 5000).requires_grad_(), meta(2, 5000, 16).requires_grad_(), meta(2, 16)
    packed = []
    with torch.autograd.graph.saved_tensors_hooks(lambda t: packed.append(t.dtype) or t, lambda t: t):
        y = LPC.apply(x, A, zi)
    assert calls == ["keep"] and torch.uint8 in packed
    torch.autograd.grad(y.sum(), (x, A))
    assert calls[-1] == ("backward", True)
    calls.clear()
    with torch.no_grad():
        y = LPC.apply(x, A, zi)
    assert calls == ["lpc"] and y.grad_fn is None
    calls.clear()
    LPC.apply(x.detach(), A.detach(), zi)
    assert calls == ["lpc"]


def test_ragged_plan_table_layout():
    """Host plan of a ragged batch: per order, offsets / lengths / first chunk ids,
    then the signal of every chunk (the layout documented in the C header)."""
    from torchlpc_amd import RaggedPlan

    lengths, orders = [100, 37, 5000, 16, 1], [2, 16, 2, 32, 2]
    p = RaggedPlan(lengths, orders, "cpu")
    assert (p.total_samples, p.total_coeffs, p.total_state) == (5154, 11306, 54)
    n, table, K, G, maxp, L, ext = p.groups[0]  # order 2: signals 0, 2, 4
    assert (n, K) == (2, 3) and L % 16 == 0
    assert ext == [5154, 11306, 54]  # ends of signal 4 in packed x, A and zi
    P = [-(-T // L) for T in (100, 5000, 1)]
    assert G == sum(P) and maxp == max(P)
    t = table.tolist()
    assert t[0:3] == [0, 137, 5153]  # sample offsets
    assert t[3:6] == [0, 792, 11304]  # coefficient offsets (T_k * N_k rows before)
    assert t[6:9] == [0, 18, 52]  # zi offsets
    assert t[9:12] == [100, 5000, 1]
    assert t[12:16] == [0, P[0], P[0] + P[1], G]
    assert t[16:] == [0] * P[0] + [1] * P[1] + [2] * P[2]
    assert [g[0] for g in p.groups] == [2, 16, 32]
    assert p.groups[1][6] == [100 + 37, 200 + 37 * 16, 2 + 16]  # order 16: signal 1 alone
    # a tensor of another dtype than the plan's is refused before any launch
    from torchlpc_amd.ragged import RaggedLPC

    f64 = torch.float64
    with pytest.raises(ValueError, match="plan was made for"):
        RaggedLPC.apply(torch.zeros(5154, dtype=f64), torch.zeros(11306, dtype=f64), torch.zeros(54, dtype=f64), p)
