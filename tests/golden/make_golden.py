"""Write the golden fixtures from the reference's own compiled CPU op.

Run in the build container (it needs /root/reference for `make -C oracle ref`):
    python tests/golden/make_golden.py

Forward outputs come straight from the reference op ``torch.ops.torchlpc.lpc`` /
``torch.ops.torchlpc.scan`` (torchlpc/csrc/scan_cpu.cpp, compiled by
oracle/Makefile into oracle/_ref/). Backward outputs come from the reference's
backward composition (torchlpc/core.py:186-229) driven through that same
compiled op: shifted coefficients, a filter pass over the reversed and padded
gradient, and dA = conj(y[t-i]) * (-dx[t]).

Each case is saved as <name>.npz (inputs and expected outputs, no pickles).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402


def make_case(rng, B, T, N, dtype):
    cplx = np.issubdtype(dtype, np.complexfloating)
    x = rng.standard_normal((B, T))
    zi = rng.standard_normal((B, N))
    g = rng.standard_normal((B, T))
    k = rng.uniform(-0.85, 0.85, (B, T // 16 + 2, N))
    kk = np.repeat(k, 16, axis=1)[:, :T]
    A = O.step_up(kk)
    if cplx:
        x = x + 1j * rng.standard_normal((B, T))
        zi = zi + 1j * rng.standard_normal((B, N))
        g = g + 1j * rng.standard_normal((B, T))
        A = A * np.exp(1j * rng.uniform(-0.3, 0.3, A.shape))
    x, A, zi, g = (v.astype(dtype) for v in (x, A, zi, g))
    tx, tA, tzi, tg = (torch.from_numpy(np.ascontiguousarray(v)) for v in (x, A, zi, g))
    y = torch.ops.torchlpc.lpc(tx, tA, tzi)
    dx, dA, dzi = O.reference_backward_torch(tg, tA, tzi, y)
    w = rng.uniform(-0.95, 0.95, (B, T)).astype(dtype)
    init = rng.standard_normal(B).astype(dtype)
    h = torch.ops.torchlpc.scan(tx, torch.from_numpy(w), torch.from_numpy(init))
    return dict(
        x=x, A=A, zi=zi, y=y.numpy(), g=g, dx=dx.numpy(), dA=dA.numpy(), dzi=dzi.numpy(),
        w=w, init=init, h=h.numpy(),
    )


CASES = [
    # name, B, T, N, dtype
    ("f32_n2_t64", 3, 64, 2, np.float32),
    ("f32_n3_t301", 3, 301, 3, np.float32),
    ("f32_n16_t301", 2, 301, 16, np.float32),
    ("f32_n1_t100", 3, 100, 1, np.float32),
    ("f64_n4_t257", 3, 257, 4, np.float64),
    ("f64_n16_t130", 2, 130, 16, np.float64),
    ("f64_n5_t3", 2, 3, 5, np.float64),
    ("c64_n2_t97", 2, 97, 2, np.complex64),
    ("c128_n3_t80", 2, 80, 3, np.complex128),
]


def main() -> None:
    O.build()
    torch.ops.load_library(str(O.REF_LIB_PATH))
    rng = np.random.default_rng(20240601)
    for name, B, T, N, dt in CASES:
        case = make_case(rng, B, T, N, dt)
        np.savez_compressed(HERE / f"{name}.npz", **case)
        print("wrote", name)


if __name__ == "__main__":
    main()
