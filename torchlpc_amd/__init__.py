"""torchlpc_amd: differentiable time-varying all-pole (LPC) filter on MI355X.

Drop-in for ``torchlpc`` (DiffAPF/torchlpc 0.8): same ``sample_wise_lpc`` and
``linear_recurrence`` API (reference ``torchlpc/__init__.py:31-104``), same
``LPC`` and ``Recurrence`` autograd Functions, computed by hand-written HIP
kernels for gfx950 behind the C ABI of ``include/torchlpc_amd.h``.

There is no CPU or Numba fallback: inputs must live on the GPU, and a missing
native library raises ``RuntimeError`` at the first call.
"""
from __future__ import annotations

from typing import Optional, Tuple, Union

import torch

from ._ext import EXTENSION_LOADED, LOAD_ERROR, hip_runtimes_mapped  # noqa: F401
from .core import LPC
from .ragged import RaggedLPC, RaggedPlan, ragged_lpc
from .recurrence import Recurrence

__all__ = [
    "sample_wise_lpc",
    "linear_recurrence",
    "LPC",
    "Recurrence",
    "RaggedPlan",
    "RaggedLPC",
    "ragged_lpc",
    "EXTENSION_LOADED",
]
__version__ = "0.8.0+mi355x"


def linear_recurrence(
    a: torch.Tensor, x: torch.Tensor, zi: Optional[torch.Tensor] = None
) -> torch.Tensor:
    """y[n] = x[n] + a[n] * y[n-1], y[-1] = zi (reference ``torchlpc/__init__.py:31-60``).

    Shapes: a (B, T), x (B, T), zi (B,). Returns (B, T).
    """
    assert a.shape == x.shape
    assert a.ndim == 2
    assert x.ndim == 2
    B, _ = a.shape
    if zi is None:
        zi = a.new_zeros(B)
    else:
        assert zi.shape == (B,)
    return Recurrence.apply(a, x, zi)


def sample_wise_lpc(
    x: torch.Tensor,
    a: torch.Tensor,
    zi: Optional[torch.Tensor] = None,
    return_zf: bool = False,
) -> Union[torch.Tensor, Tuple[torch.Tensor, torch.Tensor]]:
    """y[t] = x[t] - sum_{i=1..N} a[t,i-1] y[t-i] (reference ``torchlpc/__init__.py:63-104``).

    Shapes: x (B, T); a (B, T, N), or (B, N) for time-invariant coefficients;
    zi (B, N) with zi[:, k] = y[-1-k] (default zero). With ``return_zf`` the
    final delay values ``y[:, -N:].flip(1)`` are returned as well (shape (B, T)
    when T < N, as in the reference).
    """
    assert x.ndim == 2
    assert a.ndim in (2, 3)
    assert x.shape[0] == a.shape[0]
    if a.ndim == 3:
        assert x.shape[1] == a.shape[1]
    B, T = x.shape
    order = a.shape[-1]
    if zi is None:
        zi = a.new_zeros(B, order)
    else:
        assert zi.shape == (B, order)

    if order == 1:
        a1 = a[..., 0] if a.ndim == 3 else a[:, :1].expand(B, T)
        y = linear_recurrence(-a1, x, zi.squeeze(1))
    else:
        y = LPC.apply(x, a, zi)

    if return_zf:
        return y, y[:, -order:].flip(1)
    return y
