"""numpy front end of the C oracle, FP64 closed forms and workload generators.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblpc_oracle.so"
REF_LIB_PATH = HERE / "_ref" / "torchlpc_ref_cpu.so"

_SUFFIX = {
    np.dtype(np.float32): "f32",
    np.dtype(np.float64): "f64",
    np.dtype(np.complex64): "c64",
    np.dtype(np.complex128): "c128",
}

_lib = None


def build() -> None:
    """Compile liblpc_oracle.so (and _ref when the reference sources exist)."""
    targets = ["all"]
    if Path("/root/reference/torchlpc/csrc/scan_cpu.cpp").exists():
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(HERE), *targets], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        p, i64 = ctypes.c_void_p, ctypes.c_int64
        for sfx in _SUFFIX.values():
            getattr(_lib, f"oracle_lpc_{sfx}").argtypes = [p, p, p, p, i64, i64, i64]
            getattr(_lib, f"oracle_scan_{sfx}").argtypes = [p, p, p, p, i64, i64]
            getattr(_lib, f"oracle_lpc_backward_{sfx}").argtypes = [p] * 7 + [i64, i64, i64]
        _lib.oracle_max_threads.restype = ctypes.c_int
        _lib.oracle_set_threads.argtypes = [ctypes.c_int]
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def lpc(x, A, zi=None) -> np.ndarray:
    """y = lpc(x[B,T], A[B,T,N], zi[B,N]) with the reference loop order."""
    x = np.asarray(x)
    dt = x.dtype
    B, T = x.shape
    N = np.asarray(A).shape[2]
    x = _c(x, dt)
    A = _c(A, dt)
    zi = np.zeros((B, N), dt) if zi is None else _c(zi, dt)
    y = np.empty((B, T), dt)
    getattr(lib(), f"oracle_lpc_{_SUFFIX[dt]}")(_ptr(x), _ptr(A), _ptr(zi), _ptr(y), B, T, N)
    return y


def scan(x, w, init) -> np.ndarray:
    """h[t] = w[t] h[t-1] + x[t], h[-1] = init (reference op torchlpc::scan)."""
    x = np.asarray(x)
    dt = x.dtype
    B, T = x.shape
    x, w, init = _c(x, dt), _c(w, dt), _c(init, dt)
    h = np.empty((B, T), dt)
    getattr(lib(), f"oracle_scan_{_SUFFIX[dt]}")(_ptr(x), _ptr(w), _ptr(init), _ptr(h), B, T)
    return h


def lpc_backward(g, A, zi, y):
    """(dx, dA, dzi) of the reference backward composition for upstream grad g."""
    g = np.asarray(g)
    dt = g.dtype
    B, T = g.shape
    N = np.asarray(A).shape[2]
    g, A, zi, y = _c(g, dt), _c(A, dt), _c(zi, dt), _c(y, dt)
    dx = np.empty((B, T), dt)
    dA = np.empty((B, T, N), dt)
    dzi = np.empty((B, N), dt)
    getattr(lib(), f"oracle_lpc_backward_{_SUFFIX[dt]}")(
        _ptr(g), _ptr(A), _ptr(zi), _ptr(y), _ptr(dx), _ptr(dA), _ptr(dzi), B, T, N
    )
    return dx, dA, dzi


def scan_backward(g, w, init, h):
    """(dx, dw, dinit) of the first-order recurrence (reference recurrence.py:92-126)."""
    g = np.asarray(g)
    w = np.asarray(w)
    B, T = g.shape
    # order-1 filter with A = -w: dA = -lambda conj(h[t-1]) = -dw
    dx, dA, dzi = lpc_backward(g, -w[:, :, None], np.asarray(init)[:, None], h)
    return dx, -dA[:, :, 0], dzi[:, 0]


def reference_backward_torch(g, A, zi, y):
    """Backward of the reference CPU path (torchlpc/core.py:186-229) on torch CPU
    tensors, driving the reference's compiled op ``torch.ops.torchlpc.lpc``
    (load ``REF_LIB_PATH`` first): shifted coefficients, one filter pass over the
    reversed, zero-padded gradient, then dA = conj(y[t-i]) * (-dx[t])."""
    import torch

    B, T, N = A.shape
    S = torch.zeros((B, T + N, N), dtype=A.dtype)
    for i in range(1, N + 1):
        S[:, N - i : N - i + T, i - 1] = A[:, :, i - 1]
    g_pad = torch.cat([torch.zeros((B, N), dtype=g.dtype), g], dim=1)
    out = torch.ops.torchlpc.lpc(g_pad.flip(1), S.flip(1).conj_physical(), torch.zeros_like(zi))
    dzi = out[:, -N:]
    dx = out[:, :-N].flip(1)
    ext = torch.cat([zi.flip(1), y[:, :-1]], dim=1)
    lag = ext.unfold(1, N, 1).flip(2)  # lag[:, t, i-1] = y[t-i]
    dA = lag.conj_physical() * -dx.unsqueeze(2)
    return dx, dA, dzi


# --------------------------------------------------------------------------
# FP64 closed forms (independent of the loop above)


def biquad_impulse(r: float, theta: float, T: int) -> np.ndarray:
    """Impulse response of 1/(1 - 2 r cos(theta) z^-1 + r^2 z^-2): r^t sin((t+1)theta)/sin(theta)."""
    t = np.arange(T, dtype=np.float64)
    return r**t * np.sin((t + 1) * theta) / math.sin(theta)


def lfilter_ti(x, a, zi=None) -> np.ndarray:
    """Time-invariant filter through scipy.signal.lfilter (x[B,T], a[B,N], zi[B,N] = y[-1-k])."""
    from scipy import signal

    x = np.asarray(x, np.float64)
    a = np.asarray(a, np.float64)
    B, T = x.shape
    N = a.shape[1]
    y = np.empty_like(x)
    for b in range(B):
        den = np.concatenate([[1.0], a[b]])
        if zi is None:
            y[b] = signal.lfilter([1.0], den, x[b])
        else:
            z = signal.lfiltic([1.0], den, y=np.asarray(zi, np.float64)[b])
            y[b] = signal.lfilter([1.0], den, x[b], zi=z)[0]
    return y


def band_matrix(A_b: np.ndarray) -> np.ndarray:
    """Dense unit lower-triangular M[T,T] with M[t, t-i] = A[t, i-1] (one signal)."""
    T, N = A_b.shape
    M = np.eye(T, dtype=A_b.dtype)
    for i in range(1, N + 1):
        idx = np.arange(i, T)
        M[idx, idx - i] = A_b[i:, i - 1]
    return M


# --------------------------------------------------------------------------
# Seeded workload generators (FP64 on the CPU, then cast), one per benchmark
# workload, named for what it is and its sizes. All use numpy's PCG64 with
# seeds 1001..1005.


def step_up(k: np.ndarray) -> np.ndarray:
    """Reflection coefficients k[..., N] -> predictor a[..., N] of A(z) = 1 + sum a_i z^-i.

    Step-up (Levinson) recursion a_i^(m) = a_i^(m-1) + k_m a_{m-i}^(m-1), a_m^(m) = k_m;
    |k_m| < 1 for all m makes A(z) minimum phase, so the all-pole filter is stable.
    """
    k = np.asarray(k, np.float64)
    N = k.shape[-1]
    a = np.zeros_like(k)
    for m in range(1, N + 1):
        km = k[..., m - 1]
        prev = a[..., : m - 1].copy()
        a[..., : m - 1] = prev + km[..., None] * prev[..., ::-1]
        a[..., m - 1] = km
    return a


def _interp_anchors(values: np.ndarray, T: int, hop: int) -> np.ndarray:
    """Linear interpolation of anchor values [B, n_anchor, ...] placed every `hop` samples."""
    n_anchor = values.shape[1]
    pos = np.arange(T, dtype=np.float64) / hop
    lo = np.clip(np.floor(pos).astype(np.int64), 0, n_anchor - 1)
    hi = np.clip(lo + 1, 0, n_anchor - 1)
    frac = (pos - lo)[None, :, None] if values.ndim == 3 else (pos - lo)[None, :]
    return values[:, lo] * (1 - frac) + values[:, hi] * frac


def make_resonators_b16_t8192_n4(B=16, T=8192, N=4, seed=1001):
    """Plain forward: 2 conjugate pole pairs per signal, (r, theta) at 32 anchors."""
    assert N == 4
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, T))
    n_anchor = 32
    r = rng.uniform(0.5, 0.95, (B, n_anchor, 2))
    th = rng.uniform(0.05 * np.pi, 0.95 * np.pi, (B, n_anchor, 2))
    pos = np.linspace(0, T - 1, n_anchor)
    t = np.arange(T, dtype=np.float64)
    ri = np.empty((B, T, 2))
    ti = np.empty((B, T, 2))
    for b in range(B):
        for p in range(2):
            ri[b, :, p] = np.interp(t, pos, r[b, :, p])
            ti[b, :, p] = np.interp(t, pos, th[b, :, p])
    # (1 - 2 r1 cos th1 z^-1 + r1^2 z^-2)(1 - 2 r2 cos th2 z^-1 + r2^2 z^-2)
    c1, d1 = -2 * ri[..., 0] * np.cos(ti[..., 0]), ri[..., 0] ** 2
    c2, d2 = -2 * ri[..., 1] * np.cos(ti[..., 1]), ri[..., 1] ** 2
    A = np.stack([c1 + c2, d1 + d2 + c1 * c2, c1 * d2 + c2 * d1, d1 * d2], axis=-1)
    zi = np.zeros((B, N))
    return x, A, zi


def make_reflection_inputs(B, T, N, hop, seed, kmax=0.9, zi_std=0.1):
    """x ~ N(0,1); A from reflection coefficients k ~ U(-kmax, kmax) at anchors every
    `hop` samples, interpolated per sample and converted by step-up; zi ~ N(0, zi_std^2);
    w ~ N(0,1) (loss = sum(y * w))."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, T))
    n_anchor = (T - 1) // hop + 2
    k = rng.uniform(-kmax, kmax, (B, n_anchor, N))
    A = step_up(_interp_anchors(k, T, hop))
    zi = rng.normal(0.0, zi_std, (B, N))
    w = rng.standard_normal((B, T))
    return x, A, zi, w


def make_reflection_b256_t48000_n16(B=256, T=48000, N=16, seed=1002):
    return make_reflection_inputs(B, T, N, 240, seed)


def make_drifting_b4_t1m_n2(B=4, T=1 << 20, seed=1003):
    """Time-varying resonators: r ~ U(0.9, 0.999) per signal, angle drifting as a
    slow sinusoid (period 2^16): theta_t = theta_c + 0.1 pi sin(2 pi t / 2^16 + phi)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, T))
    r = rng.uniform(0.9, 0.999, B)
    th_c = rng.uniform(0.2 * np.pi, 0.8 * np.pi, B)
    phi = rng.uniform(0, 2 * np.pi, B)
    t = np.arange(T, dtype=np.float64)
    th = th_c[:, None] + 0.1 * np.pi * np.sin(2 * np.pi * t[None, :] / 65536 + phi[:, None])
    A = np.stack([-2 * r[:, None] * np.cos(th), np.broadcast_to((r**2)[:, None], (B, T))], axis=-1)
    zi = rng.standard_normal((B, 2))
    w = rng.standard_normal((B, T))
    return x, A, zi, w


def make_ragged_512(count=512, seed=1004, lmin=1024, lmax=65536, orders=(2, 4, 8, 16, 32)):
    """Ragged batch: lengths ~ U{lmin..lmax}, orders ~ U(orders); per-signal arrays.

    Reflection coefficients k ~ U(-0.9, 0.9) (as for B=256, T=48000) for orders up
    to 16 and U(-0.5, 0.5) at order 32: with 0.9 an order-32 time-varying filter
    grows to 1e28..1e116 over these lengths, past the float32 range of any
    implementation."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lmin, lmax + 1, count)
    ords = rng.choice(np.asarray(orders), count)
    sigs = []
    for k in range(count):
        T, N = int(lens[k]), int(ords[k])
        x = rng.standard_normal((1, T))
        n_anchor = (T - 1) // 256 + 2
        kmax = 0.9 if N <= 16 else 0.5
        kk = rng.uniform(-kmax, kmax, (1, n_anchor, N))
        A = step_up(_interp_anchors(kk, T, 256))
        zi = rng.normal(0.0, 0.1, (1, N))
        w = rng.standard_normal((1, T))
        sigs.append((x[0], A[0], zi[0], w[0]))
    return sigs


def make_ti_b1024_t16384_n32(B=1024, T=16384, N=32, seed=1005):
    """Time-invariant: one stable coefficient row per signal from k ~ U(-0.8, 0.8)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, T))
    a = step_up(rng.uniform(-0.8, 0.8, (B, N)))
    zi = rng.normal(0.0, 0.1, (B, N))
    w = rng.standard_normal((B, T))
    return x, a, zi, w


def rel_err(out, ref) -> float:
    """||out - ref||_2 / ||ref||_2 in FP64."""
    out = np.asarray(out, dtype=np.complex128 if np.iscomplexobj(out) else np.float64)
    ref = np.asarray(ref, dtype=np.complex128 if np.iscomplexobj(ref) else np.float64)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((out - ref).ravel())
    return float(num / den) if den > 0 else float(num)


if __name__ == "__main__":  # pragma: no cover
    build()
    print("oracle built:", LIB_PATH, "threads:", max_threads(), "ref:", REF_LIB_PATH.exists(), os.getcwd())
