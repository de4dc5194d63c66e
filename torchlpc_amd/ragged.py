"""Ragged batches: many signals with their own lengths and orders in one call.

The reference filters rectangular ``[B, T]`` batches with one order
(``torchlpc/__init__.py:63-104``). Here K signals are packed back to back:

* ``x``:  ``(sum T_k,)``,
* ``a``:  ``(sum T_k * N_k,)``, signal k's coefficients as ``[T_k, N_k]`` rows,
* ``zi``: ``(sum N_k,)``, signal k's initial state as in ``sample_wise_lpc``.

Each output equals ``sample_wise_lpc`` on the signal alone. With gradients on,
the float32 order-8, order-16 and order-32 groups keep their forward's workspace (saved
for backward) and their backward reads its chunk transition matrices instead of
forming them again. A :class:`RaggedPlan`
holds the metadata (lengths, orders, offsets, and per order a device table that
maps chunks to signals), built once on the host (``RaggedPlan.tune`` picks each
group's chunk length from timing); the filter then runs as one set of
chunked-scan launches per order present (summaries, carries, reruns),
forward and fused backward, so a whole ragged step can be captured in one HIP
graph. The order groups write disjoint parts of the outputs; by default they
run one after another on the caller's stream, since each group's launches fill
the chip on their own and concurrent groups slowed one another (the ragged
workload's step: 1.98 ms serial, 2.19-2.25 ms on four streams, eager or in a
graph). ``RaggedPlan(..., streams=k)`` puts them on k side streams instead.
"""
from __future__ import annotations

import threading
from typing import Any, List, Optional, Sequence, Tuple

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from ._ext import ops


class RaggedPlan:
    """Host-side layout of a ragged batch and its per-order device tables."""

    def __init__(
        self,
        lengths: Sequence[int],
        orders: Sequence[int],
        device: torch.device | str,
        dtype: torch.dtype = torch.float32,
        chunk_len: int = 0,
        streams: int = 1,
    ) -> None:
        lengths_t = torch.as_tensor(list(lengths), dtype=torch.int64)
        orders_t = torch.as_tensor(list(orders), dtype=torch.int64)
        assert lengths_t.dim() == 1 and lengths_t.shape == orders_t.shape
        assert bool((lengths_t >= 1).all()) and bool((orders_t >= 1).all())
        self.lengths = lengths_t
        self.orders = orders_t
        self.device = torch.device(device)
        self.dtype = dtype
        zero = torch.zeros(1, dtype=torch.int64)
        self.x_off = torch.cat([zero, lengths_t.cumsum(0)])
        self.a_off = torch.cat([zero, (lengths_t * orders_t).cumsum(0)])
        self.z_off = torch.cat([zero, orders_t.cumsum(0)])
        self.total_samples = int(self.x_off[-1])
        self.total_coeffs = int(self.a_off[-1])
        self.total_state = int(self.z_off[-1])
        # per order: (N, device table, K, G, max chunks, chunk length, extents);
        # extents = elements of packed x, A and zi the group's signals reach,
        # which the launches check the packed tensors against
        self.groups: List[Tuple[int, torch.Tensor, int, int, int, int, List[int]]] = []
        for n in sorted(set(orders_t.tolist())):
            assert n <= 32, "ragged batches support orders up to 32"
            self.groups.append(self._group_for(n, chunk_len))
        # streams for the groups (streams > 1): longest first onto the least loaded
        # stream, a group's cost taken as its padded order times its samples
        self._streams: Optional[List[torch.cuda.Stream]] = None
        nstreams = max(1, min(len(self.groups), streams)) if self.device.type == "cuda" else 1
        load = [0] * nstreams
        self.lanes: List[int] = [0] * len(self.groups)
        cost = []
        for gi, grp in enumerate(self.groups):
            n = grp[0]
            np_ = 1 << (n - 1).bit_length()
            cost.append((np_ * int(lengths_t[orders_t == n].sum()), gi))
        for c, gi in sorted(cost, reverse=True):
            s = min(range(nstreams), key=lambda i: load[i])
            self.lanes[gi] = s
            load[s] += c

    def _group_for(self, n: int, chunk_len: int):
        """The (N, table, K, G, max chunks, chunk length, extents) group of the
        signals of order n, chunked at chunk_len (0: the library heuristic)."""
        idx = (self.orders == n).nonzero().flatten()
        table, G, maxp, L, xe, ae, ze = ops().lpc_ragged_plan(
            self.x_off[idx].contiguous(),
            self.a_off[idx].contiguous(),
            self.z_off[idx].contiguous(),
            self.lengths[idx].contiguous(),
            n,
            self.dtype,
            chunk_len,
        )
        return (n, table.to(self.device), int(idx.numel()), G, maxp, L, [xe, ae, ze])

    def tune(
        self,
        x: torch.Tensor,
        a: torch.Tensor,
        zi: torch.Tensor,
        grad: Optional[torch.Tensor] = None,
        factors: Sequence[float] = (0.5, 0.75, 1.0, 1.5, 2.0, 3.0),
        reps: int = 3,
    ) -> dict:
        """Chunk length per order group from timing, on the caller's packed tensors.

        For every order group, the group's step (the forward and, with ``grad``,
        the fused backward) is timed with HIP events at chunk lengths around the
        library heuristic's (``factors`` times it, rounded to the staging block of
        the order), and the fastest becomes the group's plan. The chunk length sets
        the rerun lanes of the group (one lane per chunk) against the length of
        its carry chains. Returns ``{N: {"choice": L, "ms": {L: ms}}}``.
        Synchronises the device: not during HIP-graph capture. The reference has
        no tuner (``torchlpc/csrc/cuda/linear_recurrence.cu:238-239``).
        """
        assert x.is_cuda, "tuning needs GPU tensors"
        y = torch.empty_like(x)
        if grad is not None:
            g = grad.contiguous()
            dx, da, dzi = torch.empty_like(g), torch.empty_like(a), torch.empty_like(zi)
        out = {}
        for gi, grp in enumerate(self.groups):
            n, L0 = grp[0], grp[5]
            np_ = 1 << (n - 1).bit_length()
            blk = 16 if np_ < 16 else np_
            cands = sorted({max(blk, -(-int(L0 * f) // blk) * blk) for f in factors} | {L0})
            ms = {}
            best = None
            for L in cands:
                cand = grp if L == L0 else self._group_for(n, L)

                def step(c=cand):
                    # as the autograd Function runs it (kept summaries where they are reused)
                    n_, table, K, G, maxp, Lc, ext = c
                    ws = None
                    if grad is not None and keeps_summaries(c, x.dtype):
                        ws = ops().lpc_ragged_keep_(x, a, zi, table, y, K, G, maxp, n_, Lc, ext)
                    else:
                        ops().lpc_ragged_(x, a, zi, table, y, K, G, maxp, n_, Lc, ext)
                    if grad is not None:
                        ops().lpc_ragged_backward_(
                            g, a, zi, y, table, dx, da, dzi, K, G, maxp, n_, Lc, True, True, True, ext, ws
                        )

                step()
                torch.cuda.synchronize()
                times = []
                for _ in range(reps):
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    step()
                    e1.record()
                    torch.cuda.synchronize()
                    times.append(e0.elapsed_time(e1))
                ms[cand[5]] = sorted(times)[len(times) // 2]
                if best is None or ms[cand[5]] < ms[best[5]]:
                    best = cand
            self.groups[gi] = best
            out[n] = {"choice": best[5], "ms": ms}
        return out

    def run(self, fn) -> None:
        """Calls ``fn(group)`` for every order group: in order on the current
        stream, or on the plan's side streams when it was made with several."""
        self.run_indexed(lambda gi, grp: fn(grp))

    def run_indexed(self, fn) -> None:
        """``run`` with ``fn(index, group)``."""
        if self.device.type != "cuda" or len(self.groups) < 2 or max(self.lanes) == 0:
            for gi, grp in enumerate(self.groups):
                fn(gi, grp)
            return
        if self._streams is None:
            self._streams = [torch.cuda.Stream(device=self.device) for _ in range(max(self.lanes) + 1)]
        cur = torch.cuda.current_stream(self.device)
        for st in self._streams:
            st.wait_stream(cur)
        for gi, grp in enumerate(self.groups):
            with torch.cuda.stream(self._streams[self.lanes[gi]]):
                fn(gi, grp)
        for st in self._streams:
            cur.wait_stream(st)

    def split(self, packed: torch.Tensor, kind: str = "x") -> List[torch.Tensor]:
        """Views of the K signals of a packed tensor (kind "x", "a" or "zi")."""
        off = {"x": self.x_off, "a": self.a_off, "zi": self.z_off}[kind]
        out = []
        for k in range(self.lengths.numel()):
            v = packed[int(off[k]) : int(off[k + 1])]
            if kind == "a":
                v = v.view(int(self.lengths[k]), int(self.orders[k]))
            out.append(v)
        return out


# forward -> setup_context hand-over of the kept workspaces (same thread, back
# to back: Function.apply runs setup_context right after forward)
_kept = threading.local()


def keeps_summaries(grp, dtype: torch.dtype) -> bool:
    """Whether an order group's forward keeps its workspace for the backward:
    float32 orders 8, 16 and 32, whose backward then reads the forward's chunk
    transition matrices instead of forming its own (the library ignores a kept
    workspace elsewhere)."""
    return dtype == torch.float32 and grp[0] in (8, 16, 32)


class RaggedLPC(Function):
    """Packed ``y`` of the ragged batch described by a :class:`RaggedPlan`."""

    @classmethod
    def apply(cls, *args, **kwargs):
        _kept.grad_mode = torch.is_grad_enabled()
        return Function.apply.__func__(cls, *args, **kwargs)

    @staticmethod
    def forward(x: torch.Tensor, a: torch.Tensor, zi: torch.Tensor, plan: RaggedPlan) -> torch.Tensor:
        for t in (x, a, zi):
            if t.dtype != plan.dtype or t.device != plan.device:
                raise ValueError(
                    f"ragged_lpc: tensors must be {plan.dtype} on {plan.device} as the plan was "
                    f"made for, got {t.dtype} on {t.device}"
                )
        y = torch.empty_like(x)
        # Function.forward runs with grad mode off; apply recorded the caller's
        caller_grad = getattr(_kept, "grad_mode", False)
        _kept.grad_mode = False
        keep = caller_grad and not torch.compiler.is_compiling() and (
            x.requires_grad or a.requires_grad or zi.requires_grad
        )
        kept: List[Optional[torch.Tensor]] = [None] * len(plan.groups)

        def one(gi, grp):
            n, table, K, G, maxp, L, ext = grp
            if keep and keeps_summaries(grp, x.dtype):
                kept[gi] = ops().lpc_ragged_keep_(x, a, zi, table, y, K, G, maxp, n, L, ext)
            else:
                ops().lpc_ragged_(x, a, zi, table, y, K, G, maxp, n, L, ext)

        plan.run_indexed(one)
        _kept.item = (y, kept) if keep else None
        return y

    @staticmethod
    def setup_context(ctx: Any, inputs: Tuple[Any, ...], output: Any) -> None:
        _, a, zi, plan = inputs
        item = getattr(_kept, "item", None)
        _kept.item = None
        ctx.plan = plan
        ctx.kept_index = []
        ws = []
        if item is not None and item[0].data_ptr() == output.data_ptr():
            # saved tensors like the others, so saved-tensor hooks see them
            for gi, w in enumerate(item[1]):
                if w is not None:
                    ctx.kept_index.append(gi)
                    ws.append(w)
        ctx.save_for_backward(a, zi, output, *ws)

    @staticmethod
    @once_differentiable
    def backward(ctx: Any, grad_y: torch.Tensor):
        a, zi, y = ctx.saved_tensors[:3]
        kept = dict(zip(ctx.kept_index, ctx.saved_tensors[3:]))
        need_x, need_a, need_zi = ctx.needs_input_grad[:3]
        g = grad_y.contiguous()
        dx = torch.empty_like(g) if need_x else g.new_empty(0)
        da = torch.empty_like(a) if need_a else a.new_empty(0)
        dzi = torch.empty_like(zi) if need_zi else zi.new_empty(0)
        if need_x or need_a or need_zi:

            def one(gi, grp):
                n, table, K, G, maxp, L, ext = grp
                ops().lpc_ragged_backward_(
                    g, a, zi, y, table, dx, da, dzi, K, G, maxp, n, L, need_x, need_a, need_zi, ext, kept.get(gi)
                )

            ctx.plan.run_indexed(one)
        return (dx if need_x else None, da if need_a else None, dzi if need_zi else None, None)


def ragged_lpc(
    x: torch.Tensor, a: torch.Tensor, plan: RaggedPlan, zi: Optional[torch.Tensor] = None
) -> torch.Tensor:
    """Packed outputs of the ragged all-pole filter (see the module docstring)."""
    assert x.dim() == 1 and x.numel() == plan.total_samples
    assert a.dim() == 1 and a.numel() == plan.total_coeffs
    if zi is None:
        zi = a.new_zeros(plan.total_state)
    assert zi.dim() == 1 and zi.numel() == plan.total_state
    return RaggedLPC.apply(x.contiguous(), a.contiguous(), zi.contiguous(), plan)
