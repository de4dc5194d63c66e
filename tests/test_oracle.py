"""Pins of the CPU oracle (CPU only, no GPU).

Every oracle function is checked against something that does not restate its
own formula: the reference's compiled op (oracle/_ref), the golden fixtures,
closed forms (biquad impulse response, scipy lfilter, 0.9**t), finite
differences and the transpose identity. Planted-mistake tests show that the
checks catch a dropped term, a sign error and an index shift.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rand_inputs(rng, B, T, N, dtype, scale=0.3):
    x = rng.standard_normal((B, T))
    A = rng.uniform(-scale, scale, (B, T, N))
    zi = rng.standard_normal((B, N))
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * rng.standard_normal((B, T))
        A = A + 1j * rng.uniform(-scale, scale, (B, T, N)) * 0.5
        zi = zi + 1j * rng.standard_normal((B, N))
    return x.astype(dtype), A.astype(dtype), zi.astype(dtype)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden(oracle, name):
    c = load(name)
    y = oracle.lpc(c["x"], c["A"], c["zi"])
    np.testing.assert_array_equal(y, c["y"])  # same loop, same order: bit-exact
    dx, dA, dzi = oracle.lpc_backward(c["g"], c["A"], c["zi"], c["y"])
    np.testing.assert_array_equal(dx, c["dx"])
    np.testing.assert_array_equal(dzi, c["dzi"])
    if np.iscomplexobj(dA):
        # the fixture's dA product ran through torch's vectorised complex multiply
        # (fused multiply-add), the oracle's through scalar C: last-bit differences
        tol = 1e-6 if dA.dtype == np.complex64 else 1e-14
        np.testing.assert_allclose(dA, c["dA"], rtol=tol, atol=tol * np.abs(c["dA"]).max())
    else:
        np.testing.assert_array_equal(dA, c["dA"])
    h = oracle.scan(c["x"], c["w"], c["init"])
    np.testing.assert_array_equal(h, c["h"])


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
@pytest.mark.parametrize("B,T,N", [(3, 257, 5), (2, 40, 1), (4, 7, 9)])
def test_oracle_bitexact_vs_reference_op(oracle, dtype, B, T, N):
    if not oracle.REF_LIB_PATH.exists():
        pytest.skip("reference CPU op not built (oracle/_ref)")
    torch.ops.load_library(str(oracle.REF_LIB_PATH))
    rng = np.random.default_rng(B * 1000 + T + N)
    x, A, zi = rand_inputs(rng, B, T, N, dtype)
    y = oracle.lpc(x, A, zi)
    yr = torch.ops.torchlpc.lpc(torch.from_numpy(x), torch.from_numpy(A), torch.from_numpy(zi))
    np.testing.assert_array_equal(y, yr.numpy())
    w = rng.uniform(-0.9, 0.9, (B, T)).astype(dtype)
    init = rng.standard_normal(B).astype(dtype)
    h = oracle.scan(x, w, init)
    hr = torch.ops.torchlpc.scan(torch.from_numpy(x), torch.from_numpy(w), torch.from_numpy(init))
    np.testing.assert_array_equal(h, hr.numpy())


def test_biquad_impulse_closed_form(oracle):
    T = 400
    for r, th in [(0.95, 0.7), (0.5, 2.5), (0.999, 0.05)]:
        A = np.zeros((1, T, 2))
        A[..., 0] = -2 * r * np.cos(th)
        A[..., 1] = r * r
        x = np.zeros((1, T))
        x[0, 0] = 1.0
        y = oracle.lpc(x, A)
        ref = oracle.biquad_impulse(r, th, T)
        assert np.max(np.abs(y[0] - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


def test_time_invariant_matches_lfilter(oracle):
    rng = np.random.default_rng(7)
    B, T, N = 3, 500, 6
    a = oracle.step_up(rng.uniform(-0.8, 0.8, (B, N)))
    x = rng.standard_normal((B, T))
    zi = rng.standard_normal((B, N))
    A = np.broadcast_to(a[:, None, :], (B, T, N))
    y = oracle.lpc(x, A, zi)
    np.testing.assert_allclose(y, oracle.lfilter_ti(x, a, zi), rtol=1e-10, atol=1e-10)


def test_scan_geometric_decay(oracle):
    # the printed demo of the reference's parallel scan: decay 0.9, impulse at t=0
    B, T = 4, 2048
    w = np.full((B, T), 0.9)
    x = np.zeros((B, T))
    x[:, 0] = 1.0
    h = oracle.scan(x, w, np.zeros(B))
    np.testing.assert_allclose(h, np.broadcast_to(0.9 ** np.arange(T), (B, T)), rtol=1e-12, atol=1e-300)


def test_linearity_and_chaining(oracle):
    rng = np.random.default_rng(11)
    B, T, N = 2, 300, 4
    x1, A, zi1 = rand_inputs(rng, B, T, N, np.float64)
    x2, _, zi2 = rand_inputs(rng, B, T, N, np.float64)
    y = oracle.lpc(2.0 * x1 - 3.0 * x2, A, 2.0 * zi1 - 3.0 * zi2)
    ref = 2.0 * oracle.lpc(x1, A, zi1) - 3.0 * oracle.lpc(x2, A, zi2)
    np.testing.assert_allclose(y, ref, rtol=1e-10, atol=1e-10)
    # zi/zf chaining: filtering the concatenation equals two chained calls
    yfull = oracle.lpc(x1, A, zi1)
    k = 123
    y_a = oracle.lpc(x1[:, :k], A[:, :k], zi1)
    zf = y_a[:, ::-1][:, :N]  # zf[k] = y[T-1-k]
    y_b = oracle.lpc(x1[:, k:], A[:, k:], zf)
    np.testing.assert_allclose(np.concatenate([y_a, y_b], 1), yfull, rtol=1e-12, atol=1e-12)


def test_triangular_solve(oracle):
    rng = np.random.default_rng(3)
    T, N = 60, 3
    x, A, _ = rand_inputs(rng, 1, T, N, np.float64)
    M = oracle.band_matrix(A[0])
    np.testing.assert_allclose(oracle.lpc(x, A)[0], np.linalg.solve(M, x[0]), rtol=1e-10, atol=1e-12)


def _fd_grads(oracle, x, A, zi, g, eps=1e-6):
    """Central finite differences of L = <g, lpc(x, A, zi)> (real FP64)."""
    def L(x_, A_, zi_):
        return float(np.sum(g * oracle.lpc(x_, A_, zi_)))

    out = []
    for arr in (x, A, zi):
        d = np.zeros_like(arr)
        it = np.nditer(arr, flags=["multi_index"])
        for _ in it:
            idx = it.multi_index
            p = [x.copy(), A.copy(), zi.copy()]
            m = [x.copy(), A.copy(), zi.copy()]
            k = [id(x), id(A), id(zi)].index(id(arr))
            p[k][idx] += eps
            m[k][idx] -= eps
            d[idx] = (L(*p) - L(*m)) / (2 * eps)
        out.append(d)
    return out


def test_backward_finite_differences(oracle):
    rng = np.random.default_rng(5)
    B, T, N = 2, 9, 3
    x, A, zi = rand_inputs(rng, B, T, N, np.float64)
    g = rng.standard_normal((B, T))
    y = oracle.lpc(x, A, zi)
    dx, dA, dzi = oracle.lpc_backward(g, A, zi, y)
    fx, fA, fzi = _fd_grads(oracle, x, A, zi, g)
    np.testing.assert_allclose(dx, fx, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(dA, fA, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(dzi, fzi, rtol=1e-6, atol=1e-8)


def test_backward_transpose_identity(oracle):
    # <M^-1 u, v> = <u, M^-T v>: the adjoint of the zero-state filter
    rng = np.random.default_rng(9)
    B, T, N = 3, 200, 5
    u, A, _ = rand_inputs(rng, B, T, N, np.float64)
    v = rng.standard_normal((B, T))
    zero = np.zeros((B, N))
    lhs = np.sum(oracle.lpc(u, A, zero) * v)
    dx, _, _ = oracle.lpc_backward(v, A, zero, oracle.lpc(u, A, zero))
    assert abs(lhs - np.sum(u * dx)) < 1e-9 * max(1.0, abs(lhs))


def test_scan_backward_matches_order1_fd(oracle):
    rng = np.random.default_rng(13)
    B, T = 2, 12
    x = rng.standard_normal((B, T))
    w = rng.uniform(-0.9, 0.9, (B, T))
    init = rng.standard_normal(B)
    g = rng.standard_normal((B, T))
    h = oracle.scan(x, w, init)
    dx, dw, dinit = oracle.scan_backward(g, w, init, h)
    eps = 1e-6
    for arr, d in ((x, dx), (w, dw), (init, dinit)):
        for idx in np.ndindex(arr.shape):
            orig = arr[idx]
            arr[idx] = orig + eps
            lp = np.sum(g * oracle.scan(x, w, init))
            arr[idx] = orig - eps
            lm = np.sum(g * oracle.scan(x, w, init))
            arr[idx] = orig
            assert abs((lp - lm) / (2 * eps) - d[idx]) < 1e-6


# ---------------------------------------------------------------------------
# Planted mistakes: a wrong oracle must fail the pins above.


def _mutant_lpc(x, A, zi, mode):
    B, T = x.shape
    N = A.shape[2]
    out = np.concatenate([zi[:, ::-1], x], axis=1).astype(np.float64)
    for b in range(B):
        for t in range(T):
            acc = out[b, t + N]
            for i in range(N):
                if mode == "drop_last" and i == N - 1:
                    continue
                lag = i + 2 if mode == "shift" else i + 1
                term = A[b, t, i] * out[b, t + N - lag] if t + N - lag >= 0 else 0.0
                acc = acc + term if mode == "sign" else acc - term
            out[b, t + N] = acc
    return out[:, N:]


@pytest.mark.parametrize("mode", ["drop_last", "sign", "shift"])
def test_planted_mistakes_are_caught(oracle, mode):
    r, th, T = 0.9, 0.8, 64
    A = np.zeros((1, T, 2))
    A[..., 0] = -2 * r * np.cos(th)
    A[..., 1] = r * r
    x = np.zeros((1, T))
    x[0, 0] = 1.0
    good = oracle.lpc(x, A, np.zeros((1, 2)))
    bad = _mutant_lpc(x, A, np.zeros((1, 2)), mode)
    ref = oracle.biquad_impulse(r, th, T)
    assert np.max(np.abs(good[0] - ref)) < 1e-9
    assert np.max(np.abs(bad[0] - ref)) > 1e-3
