"""Helpers shared by the GPU parity tests (inputs, tolerances, error metric)."""
from __future__ import annotations

import numpy as np
import torch

from oracle import oracle as O

# Relative L2 error bars against the FP64 oracle (the stated per-result bars).
REL_TOL = {
    np.float32: 1e-5,
    np.float64: 1e-12,
    np.complex64: 1e-5,
    np.complex128: 1e-12,
}
HI = {np.float32: np.float64, np.float64: np.float64, np.complex64: np.complex128, np.complex128: np.complex128}


def stable_inputs(rng, B, T, N, dtype, hop=64, kmax=None):
    """x ~ N(0,1), zi ~ N(0,1), A from reflection coefficients k ~ U(-kmax, kmax) at
    anchors every `hop` samples, linearly interpolated per sample and converted by
    the step-up recursion. The default kmax = min(0.8, 4/N) keeps the time-varying
    filter well conditioned at every order (the FP32 reference loop stays below
    2e-7 relative error on these inputs), so the bars test the kernels, not the
    conditioning of the input."""
    cplx = np.issubdtype(dtype, np.complexfloating)
    if kmax is None:
        kmax = min(0.8, 4.0 / N)
    x = rng.standard_normal((B, T))
    zi = rng.standard_normal((B, N))
    nk = (T - 1) // hop + 2
    k = O._interp_anchors(rng.uniform(-kmax, kmax, (B, nk, N)), T, hop)
    A = O.step_up(k)
    if cplx:
        x = x + 1j * rng.standard_normal((B, T))
        zi = zi + 1j * rng.standard_normal((B, N))
        A = A * np.exp(1j * rng.uniform(-0.2, 0.2, A.shape))
    return x.astype(dtype), A.astype(dtype), zi.astype(dtype)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def rel(out, ref):
    return O.rel_err(out, ref)


def up(a, dtype):
    return np.asarray(a).astype(HI[dtype])
