"""First-order recurrence ``Recurrence`` Function (drop-in for ``torchlpc.recurrence``).

Reference: ``torchlpc/recurrence.py:73-168``. ``Recurrence.apply(decay, impulse,
initial_state)`` computes h[t] = decay[t] * h[t-1] + impulse[t], h[-1] =
initial_state. Forward and the fused adjoint run in the HIP library
(``torch.ops.torchlpc_amd.scan`` / ``scan_backward``, the order-1 case of the
chunked filter); under ``create_graph`` the backward is written with
``Recurrence.apply`` so it stays differentiable.
"""
from __future__ import annotations

from typing import Any, List, Optional, Tuple

import torch
import torch.nn.functional as F
from torch.autograd import Function

from . import autotune
from ._ext import ops


def composite_backward(grad_out, decay, init, out, need_decay, need_impulse, need_init):
    """mu[t] = g[t] + conj(decay[t+1]) mu[t+1]; differentiable (create_graph path)."""
    T = grad_out.shape[1]
    d_rev = F.pad(decay, (0, 1)).flip(1).conj()  # d_rev[u] = conj(decay[T-u]), d_rev[0] = 0
    g_rev = F.pad(grad_out, (1, 0)).flip(1)  # g_rev[u] = g[T-1-u], g_rev[T] = 0
    mu_rev = Recurrence.apply(d_rev, g_rev, torch.zeros_like(init))
    d_init = mu_rev[:, T] if need_init else None
    mu = mu_rev[:, :T].flip(1)
    d_impulse = mu if need_impulse else None
    d_decay = None
    if need_decay:
        h_prev = torch.cat([init[:, None], out[:, :-1]], dim=1)
        d_decay = h_prev.conj() * mu
    return d_decay, d_impulse, d_init


class Recurrence(Function):
    @staticmethod
    def forward(decay: torch.Tensor, impulse: torch.Tensor, initial_state: torch.Tensor) -> torch.Tensor:
        _, chunk = autotune.lookup("scan", impulse, decay)
        return ops().scan(impulse, decay, initial_state, chunk)

    @staticmethod
    def setup_context(ctx: Any, inputs: List[Any], output: Any) -> Any:
        decay, _, initial_state = inputs
        ctx.save_for_backward(decay, initial_state, output)
        ctx.save_for_forward(decay, initial_state, output)

    @staticmethod
    def backward(
        ctx: Any, grad_out: torch.Tensor
    ) -> Tuple[Optional[torch.Tensor], Optional[torch.Tensor], Optional[torch.Tensor]]:
        decay, init, out = ctx.saved_tensors
        need_decay, need_impulse, need_init = ctx.needs_input_grad[:3]
        if torch.is_grad_enabled():
            return composite_backward(grad_out, decay, init, out, need_decay, need_impulse, need_init)
        _, chunk = autotune.lookup("scan", grad_out, decay)
        dx, dw, dinit = ops().scan_backward(
            grad_out, decay, init, out, need_impulse, need_decay, need_init, chunk
        )
        return (
            dw if need_decay else None,
            dx if need_impulse else None,
            dinit if need_init else None,
        )

    @staticmethod
    def jvp(
        ctx: Any,
        grad_decay: Optional[torch.Tensor],
        grad_impulse: Optional[torch.Tensor],
        grad_initial_state: Optional[torch.Tensor],
    ) -> torch.Tensor:
        decay, init, out = ctx.saved_tensors
        t_init = grad_initial_state if grad_initial_state is not None else torch.zeros_like(init)
        t_imp = grad_impulse if grad_impulse is not None else torch.zeros_like(out)
        if grad_decay is not None:
            h_prev = torch.cat([init[:, None], out[:, :-1]], dim=1)
            t_imp = t_imp + h_prev * grad_decay
        return Recurrence.apply(decay, t_imp, t_init)

    @staticmethod
    def vmap(info, in_dims, decay, impulse, initial_state):
        def front(t, d):
            return t.expand(info.batch_size, *t.shape) if d is None else t.movedim(d, 0)

        args = [front(t, d) for t, d in zip((decay, impulse, initial_state), in_dims)]
        flat = [t.reshape(-1, *t.shape[2:]) for t in args]
        out = Recurrence.apply(*flat)
        return out.reshape(info.batch_size, -1, *out.shape[1:]), 0
