"""``LPC`` autograd Function: the drop-in for ``torchlpc.core.LPC``.

Reference: ``torchlpc/core.py:159-265``. Same ``apply(x, A, zi)`` signature,
``setup_context`` saving ``(A, zi, y)``, ``backward``, forward-mode ``jvp`` and a
custom ``vmap``. The arithmetic runs in the HIP kernels of
``libtorchlpc_amd_hip.so``:

* forward: ``torch.ops.torchlpc_amd.lpc`` (chunked scan or per-signal kernel);
* backward: ``torch.ops.torchlpc_amd.lpc_backward``, the fused adjoint that gives
  dx, dA and dzi in one reverse-time pass. A forward whose inputs require grad
  runs ``lpc_keep`` instead, which also returns its workspace: the chunk
  transition matrices in it depend on A only, and the backward reuses them (the
  adjoint's are their transposes). When the backward itself is being
  differentiated (``create_graph=True``), the gradients are instead expressed
  through ``LPC.apply`` (one more forward filter on the reversed, shifted
  coefficients, as the reference does at ``core.py:194-227``) so that they stay
  differentiable.

Long signals (``T >= RECOMPUTE_MIN_T``, real data, order up to 4): the forward
runs ``lpc_carries`` and the Function saves x and the forward's chunk carries
instead of y; the backward (``lpc_backward_recompute``) rebuilds y chunk by
chunk inside its adjoint pass. Between the passes that keeps B*P*N float64
carries (P = T / chunk length) instead of B*T samples of y; the reference saves
y (``torchlpc/core.py:179-184``).

Besides ``A[B,T,N]`` the Function accepts a time-invariant ``a[B,N]`` (its
gradient is then ``[B,N]``); a time-broadcast ``A`` (``A.stride(1) == 0``) is
read without being materialised.
"""
from __future__ import annotations

import threading
from typing import Any, List, Optional, Tuple

import torch
import torch.nn.functional as F
from torch.autograd import Function

from . import autotune
from ._ext import ops


def lagged_outputs(y: torch.Tensor, zi: torch.Tensor, order: int) -> torch.Tensor:
    """Y[b, t, i-1] = y[b, t-i] for i = 1..order, with y[b, -k] = zi[b, k-1]."""
    T = y.shape[1]
    ext = torch.cat([zi.flip(1), y], dim=1)  # ext[:, order + u] = y_u
    t = torch.arange(T, device=y.device)[:, None]
    i = torch.arange(1, order + 1, device=y.device)[None, :]
    return ext[:, t - i + order]


def shifted_coefficients(A: torch.Tensor) -> torch.Tensor:
    """S[b, s, i-1] = A[b, s-N+i, i-1] for s in [0, T+N) (zero outside [0, T)).

    Row s of S holds the coefficients that multiply lambda[s-N+i] in the adjoint
    recurrence written in direct form (reference: the flip/pad/reshape at
    ``torchlpc/core.py:194-201``).
    """
    N = A.shape[2]
    cols = [F.pad(A[:, :, i - 1], (N - i, i)) for i in range(1, N + 1)]
    return torch.stack(cols, dim=2)


def _as_3d(A: torch.Tensor, T: int) -> torch.Tensor:
    return A[:, None, :].expand(A.shape[0], T, A.shape[1]) if A.dim() == 2 else A


def composite_backward(grad_y, A, zi, y, need_x, need_A, need_zi):
    """Differentiable adjoint built from ``LPC.apply`` (used under create_graph)."""
    B, T = grad_y.shape
    A3 = _as_3d(A, T)
    N = A3.shape[2]
    S = shifted_coefficients(A3)  # [B, T+N, N]
    g_ext = F.pad(grad_y, (N, 0))  # zero gradient at positions -N..-1
    lam_rev = LPC.apply(g_ext.flip(1), S.flip(1).conj(), torch.zeros_like(zi))
    dzi = lam_rev[:, T:] if need_zi else None
    lam = lam_rev[:, :T].flip(1)
    dx = lam if need_x else None
    dA = None
    if need_A:
        dA = -(lagged_outputs(y, zi, N).conj() * lam.unsqueeze(2))
        if A.dim() == 2:
            dA = dA.sum(1)
    return dx, dA, dzi


# forward -> setup_context hand-over of the kept workspace (same thread, back to
# back: Function.apply runs setup_context right after forward)
_kept = threading.local()


def _keep_summaries(x: torch.Tensor, A: torch.Tensor, zi: torch.Tensor) -> bool:
    """Whether the forward keeps its workspace for the backward: only when a
    backward can follow (the caller's grad mode was on and an input requires
    grad) and not while torch.compile traces (the fake kernels then see the
    plain op). Function.forward itself runs with grad mode off, so LPC.apply
    records the caller's mode first."""
    caller_grad = getattr(_kept, "grad_mode", False)
    _kept.grad_mode = False
    if torch.compiler.is_compiling() or not caller_grad:
        return False
    return x.requires_grad or A.requires_grad or zi.requires_grad


# Signals at least this long keep the forward's chunk carries instead of y when
# the backward can rebuild y from them (set to a huge value to always save y).
RECOMPUTE_MIN_T = 1 << 16


def _recompute_plan(x: torch.Tensor, A: torch.Tensor, plan: Tuple[int, int]) -> Optional[Tuple[int, int]]:
    """The (strategy, chunk length) of a y-rebuilding backward for this forward, or
    None: long time-varying real signals of order up to 4 whose chunks the
    rebuilding rerun takes (at most 256 samples, so a plan with longer chunks
    is replaced by 256-sample chunks)."""
    from ._ext import recompute_supported

    if x.shape[1] < RECOMPUTE_MIN_T or A.dim() != 3 or A.shape[-1] > 4 or x.dtype not in (torch.float32, torch.float64):
        return None
    for cand in (plan, (1, 256)):
        if recompute_supported(x, A, *cand):
            return cand
    return None


class LPC(Function):
    """y[t] = x[t] - sum_{i=1..N} A[t,i] y[t-i], with y[-k] = zi[k-1]."""

    @classmethod
    def apply(cls, *args, **kwargs):
        _kept.grad_mode = torch.is_grad_enabled()
        return Function.apply.__func__(cls, *args, **kwargs)

    @staticmethod
    def forward(x: torch.Tensor, A: torch.Tensor, zi: torch.Tensor) -> torch.Tensor:
        strategy, chunk = autotune.lookup("lpc_forward", x, A)
        if _keep_summaries(x, A, zi):
            rc = _recompute_plan(x, A, (strategy, chunk))
            if rc is not None:
                # keep the chunk carries; the backward rebuilds y from them
                y, ws = ops().lpc_carries(x, A, zi, *rc)
                _kept.item = (y, ws, rc[0], rc[1], True)
                return y
            # the backward runs with the forward's strategy and chunk length, so
            # that the kept summaries match its chunking
            y, ws = ops().lpc_keep(x, A, zi, strategy, chunk)
            _kept.item = (y, ws, strategy, chunk, False) if ws.numel() > 0 else None
            return y
        return ops().lpc(x, A, zi, strategy, chunk)

    @staticmethod
    def setup_context(ctx: Any, inputs: List[Any], output: Any) -> Any:
        x, A, zi = inputs
        kept = getattr(_kept, "item", None)
        _kept.item = None
        ctx.kept_plan = None
        ctx.recompute = False
        if kept is not None and kept[0].data_ptr() == output.data_ptr():
            # the workspace is a saved tensor like the others, so saved-tensor
            # hooks (offloading, checkpointing) see it
            ctx.kept_plan = kept[2:4]
            ctx.recompute = kept[4]
            if ctx.recompute:
                ctx.save_for_backward(A, zi, x, kept[1])
            else:
                ctx.save_for_backward(A, zi, output, kept[1])
        else:
            ctx.save_for_backward(A, zi, output)
        ctx.save_for_forward(A, zi, output)

    @staticmethod
    def backward(
        ctx: Any, grad_y: torch.Tensor
    ) -> Tuple[Optional[torch.Tensor], Optional[torch.Tensor], Optional[torch.Tensor]]:
        need_x, need_A, need_zi = ctx.needs_input_grad[:3]
        if getattr(ctx, "recompute", False):
            A, zi, x, ws = ctx.saved_tensors
            strategy, chunk = ctx.kept_plan
            if torch.is_grad_enabled():
                y = ops().lpc(x, A, zi, strategy, chunk)
                return composite_backward(grad_y, A, zi, y, need_x, need_A, need_zi)
            dx, dA, dzi = ops().lpc_backward_recompute(
                grad_y.contiguous(), A, zi, x, ws, need_x, need_A, need_zi, strategy, chunk
            )
            return (dx if need_x else None, dA if need_A else None, dzi if need_zi else None)
        A, zi, y = ctx.saved_tensors[:3]
        if torch.is_grad_enabled():
            return composite_backward(grad_y, A, zi, y, need_x, need_A, need_zi)
        kept_plan = getattr(ctx, "kept_plan", None)
        if kept_plan is not None:
            ws = ctx.saved_tensors[3]
            strategy, chunk = kept_plan
            dx, dA, dzi = ops().lpc_backward(
                grad_y, A, zi, y, need_x, need_A, need_zi, strategy, chunk, ws
            )
        else:
            strategy, chunk = autotune.lookup("lpc_backward", grad_y, A)
            dx, dA, dzi = ops().lpc_backward(
                grad_y, A, zi, y, need_x, need_A, need_zi, strategy, chunk
            )
        return (dx if need_x else None, dA if need_A else None, dzi if need_zi else None)

    @staticmethod
    def jvp(
        ctx: Any,
        grad_x: Optional[torch.Tensor],
        grad_A: Optional[torch.Tensor],
        grad_zi: Optional[torch.Tensor],
    ) -> torch.Tensor:
        A, zi, y = ctx.saved_tensors[:3]
        N = A.shape[-1]
        tx = grad_x if grad_x is not None else torch.zeros_like(y)
        tzi = grad_zi if grad_zi is not None else torch.zeros_like(zi)
        if grad_A is not None:
            gA = _as_3d(grad_A, y.shape[1])
            tx = tx - (lagged_outputs(y, zi, N) * gA).sum(2)
        return LPC.apply(tx, A, tzi)

    @staticmethod
    def vmap(info, in_dims, x, A, zi):
        def front(t, d):
            return t.expand(info.batch_size, *t.shape) if d is None else t.movedim(d, 0)

        args = [front(t, d) for t, d in zip((x, A, zi), in_dims)]
        flat = [t.reshape(-1, *t.shape[2:]) for t in args]
        y = LPC.apply(*flat)
        return y.reshape(info.batch_size, -1, *y.shape[1:]), 0
