"""Ragged batches (packed signals with their own lengths and orders) against the
FP64 oracle run on each signal alone, including the full 512-signal workload
(lengths ~ U{1024..65536}, orders in {2, 4, 8, 16, 32}) and a HIP-graph capture
of its forward + backward."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.gpu_util import REL_TOL, dev, host, rel

pytestmark = pytest.mark.gpu


def _pack(sigs, dtype):
    x = np.concatenate([s[0] for s in sigs]).astype(dtype)
    a = np.concatenate([s[1].ravel() for s in sigs]).astype(dtype)
    zi = np.concatenate([s[2] for s in sigs]).astype(dtype)
    g = np.concatenate([s[3] for s in sigs]).astype(dtype)
    return x, a, zi, g


def _check(sigs, dtype, plan, y, dx, da, dzi, tol):
    xs = plan.split(torch.from_numpy(np.arange(plan.total_samples)), "x")
    worst = 0.0
    yo, dxo, dao, dzo = host(y), host(dx), host(da), host(dzi)
    for k, (x, A, zi, g) in enumerate(sigs):
        up = lambda v: np.asarray(v).astype(dtype).astype(np.complex128 if np.iscomplexobj(v) else np.float64)  # noqa: E731
        X, AA, Z, G = up(x)[None], up(A)[None], up(zi)[None], up(g)[None]
        y64 = O.lpc(X, AA, Z)
        rx, rA, rzi = O.lpc_backward(G, AA, Z, y64)
        s0, s1 = int(plan.x_off[k]), int(plan.x_off[k + 1])
        a0, a1 = int(plan.a_off[k]), int(plan.a_off[k + 1])
        z0, z1 = int(plan.z_off[k]), int(plan.z_off[k + 1])
        errs = [
            rel(yo[s0:s1], y64[0]),
            rel(dxo[s0:s1], rx[0]),
            rel(dao[a0:a1], rA[0].ravel()),
            rel(dzo[z0:z1], rzi[0]),
        ]
        worst = max(worst, *errs)
        assert max(errs) <= tol, (k, len(x), A.shape[1], errs)
    del xs
    return worst


def _mixed(rng, lengths, orders, cplx=False):
    sigs = []
    for T, N in zip(lengths, orders):
        x = rng.standard_normal(T)
        kmax = min(0.8, 4.0 / N)
        k = O._interp_anchors(rng.uniform(-kmax, kmax, (1, (T - 1) // 64 + 2, N)), T, 64)
        A = O.step_up(k)[0]
        zi = rng.standard_normal(N)
        g = rng.standard_normal(T)
        if cplx:
            x = x + 1j * rng.standard_normal(T)
            A = A * np.exp(1j * rng.uniform(-0.2, 0.2, A.shape))
            zi = zi + 1j * rng.standard_normal(N)
            g = g + 1j * rng.standard_normal(T)
        sigs.append((x, A, zi, g))
    return sigs


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64])
def test_ragged_mixed_lengths_and_orders(dtype):
    from torchlpc_amd import RaggedPlan, ragged_lpc

    rng = np.random.default_rng(11)
    lengths = [1, 15, 16, 17, 33, 100, 255, 256, 257, 1000, 3001, 4096, 40, 7, 2048, 999]
    orders = [1, 2, 3, 4, 8, 16, 17, 32, 2, 16, 32, 4, 1, 16, 8, 3]
    sigs = _mixed(rng, lengths, orders, np.issubdtype(dtype, np.complexfloating))
    x, a, zi, g = _pack(sigs, dtype)
    X, AA, Z = (dev(v).requires_grad_() for v in (x, a, zi))
    plan = RaggedPlan(lengths, orders, X.device, X.dtype)
    y = ragged_lpc(X, AA, plan, Z)
    y.backward(dev(g))
    _check(sigs, dtype, plan, y, X.grad, AA.grad, Z.grad, REL_TOL[dtype])


@pytest.mark.parametrize("streams", [1, 4])
def test_ragged_full_workload_and_graph_replay(streams):
    """The order groups in order on one stream (the default), or on side streams
    (parallel branches of the captured graph)."""
    from torchlpc_amd import RaggedPlan
    from torchlpc_amd.ragged import RaggedLPC

    sigs = O.make_ragged_512()
    lengths = [len(s[0]) for s in sigs]
    orders = [s[1].shape[1] for s in sigs]
    x, a, zi, g = _pack(sigs, np.float32)
    X, AA, Z, Gd = dev(x), dev(a), dev(zi), dev(g)
    plan = RaggedPlan(lengths, orders, X.device, X.dtype, streams=streams)
    assert max(plan.lanes) == min(streams, 5) - 1
    ops = torch.ops.torchlpc_amd

    def step():
        y = RaggedLPC.forward(X, AA, Z, plan)
        dx, da, dzi = torch.empty_like(X), torch.empty_like(AA), torch.empty_like(Z)

        def bwd(grp):
            n, table, K, G, maxp, L, ext = grp
            ops.lpc_ragged_backward_(Gd, AA, Z, y, table, dx, da, dzi, K, G, maxp, n, L, True, True, True, ext)

        plan.run(bwd)
        return y, dx, da, dzi

    eager = step()
    worst = _check(sigs, np.float32, plan, *eager, 1e-5)
    print("ragged workload worst rel err", worst)
    eager = [t.clone() for t in eager]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = step()
    for r in range(3):
        torch.randn(1 << 20, device="cuda").mul_(3.0)
        graph.replay()
        torch.cuda.synchronize()
        for e, o in zip(eager, outs):
            assert torch.equal(e, o), r


def test_ragged_ops_refuse_tensors_smaller_than_the_plan():
    """The raw ops check the packed tensors against the extents of the plan's
    signals on the host, before anything is launched."""
    from torchlpc_amd import RaggedPlan

    lengths, orders = [300, 1000, 40], [4, 16, 4]
    dev0 = torch.device("cuda:0")
    plan = RaggedPlan(lengths, orders, dev0, torch.float32)
    x = torch.zeros(plan.total_samples, device=dev0)
    a = torch.zeros(plan.total_coeffs, device=dev0)
    zi = torch.zeros(plan.total_state, device=dev0)
    ops = torch.ops.torchlpc_amd
    n, table, K, G, maxp, L, ext = plan.groups[1]  # order 16: signal 1 ends at sample 1300
    y = torch.empty_like(x)
    ops.lpc_ragged_(x, a, zi, table, y, K, G, maxp, n, L, ext)  # whole tensors pass
    short = x[:1299]
    with pytest.raises(RuntimeError, match="smaller than the plan"):
        ops.lpc_ragged_(short, a, zi, table, torch.empty_like(short), K, G, maxp, n, L, ext)
    with pytest.raises(RuntimeError, match="smaller than the plan"):
        ops.lpc_ragged_(x, a[: ext[1] - 1], zi, table, y, K, G, maxp, n, L, ext)
    g = torch.zeros_like(x)
    dx, da, dzi = torch.empty_like(x), torch.empty_like(a), torch.empty_like(zi)
    with pytest.raises(RuntimeError, match="smaller than the plan"):
        zs = zi[: ext[2] - 1]
        ops.lpc_ragged_backward_(g, a, zs, y, table, dx, da, torch.empty_like(zs), K, G, maxp, n, L,
                                 True, True, True, ext)
    torch.cuda.synchronize()
    assert torch.equal(y[300:1300], torch.zeros(1000, device=dev0))  # zero input, zero state
    del dx, da, dzi


def test_ragged_tune_per_group_chunk_lengths():
    """RaggedPlan.tune picks a chunk length per order group from timing; the
    filter then runs with the chosen plans and stays within the tolerance of the
    FP64 oracle (a chunk boundary changes only float64 rounding)."""
    from torchlpc_amd import RaggedPlan, ragged_lpc

    rng = np.random.default_rng(23)
    lengths = [3000, 5000, 700, 4096, 2500, 6000, 1200, 3333]
    orders = [2, 4, 8, 16, 32, 16, 8, 32]
    sigs = _mixed(rng, lengths, orders)
    x, a, zi, g = _pack(sigs, np.float32)
    X, AA, Z, Gd = dev(x), dev(a), dev(zi), dev(g)
    plan = RaggedPlan(lengths, orders, X.device, X.dtype)
    before = {grp[0]: grp[5] for grp in plan.groups}
    res = plan.tune(X, AA, Z, Gd, reps=2)
    assert set(res) == set(before)
    for grp in plan.groups:
        n, L = grp[0], grp[5]
        assert L == res[n]["choice"] and L in res[n]["ms"] and before[n] in res[n]["ms"]
        assert L % (16 if n <= 16 else 32) == 0
    Xg, Ag, Zg = (t.clone().requires_grad_() for t in (X, AA, Z))
    y = ragged_lpc(Xg, Ag, plan, Zg)
    y.backward(Gd)
    _check(sigs, np.float32, plan, y, Xg.grad, Ag.grad, Zg.grad, REL_TOL[np.float32])


def test_ragged_kept_summaries_backward():
    """With gradients on, the float32 order-8, order-16 and order-32 groups keep their
    forward's workspace and the backward reads the forward's chunk transition
    matrices (transposed) instead of forming its own: the gradients match the
    FP64 oracle and the backward that forms its own summaries, including chunks
    that end past a signal's end, one-chunk signals, rows that are not 16-byte
    aligned (an odd-length order-1 signal before them), a repeated backward on
    the same workspace, and a HIP-graph replay of the kept step."""
    from torchlpc_amd import RaggedPlan, ragged_lpc
    from torchlpc_amd.ragged import keeps_summaries

    rng = np.random.default_rng(29)
    lengths = [5, 3001, 4096, 100, 2500, 6000, 777, 3333, 64, 9000]
    orders = [1, 16, 32, 16, 32, 16, 32, 8, 32, 16]
    sigs = _mixed(rng, lengths, orders)
    x, a, zi, g = _pack(sigs, np.float32)
    X, AA, Z, Gd = dev(x), dev(a), dev(zi), dev(g)
    plan = RaggedPlan(lengths, orders, X.device, X.dtype, chunk_len=128)
    assert any(keeps_summaries(grp, X.dtype) for grp in plan.groups)
    ops = torch.ops.torchlpc_amd

    # autograd: the kept path, checked against FP64
    Xg, Ag, Zg = (t.clone().requires_grad_() for t in (X, AA, Z))
    y = ragged_lpc(Xg, Ag, plan, Zg)
    y.backward(Gd)
    _check(sigs, np.float32, plan, y, Xg.grad, Ag.grad, Zg.grad, REL_TOL[np.float32])

    # raw ops: kept against formed summaries, and the kept workspace used twice
    def run(kept_mode, steps=1):
        y = torch.empty_like(X)
        ws = {}
        for gi, grp in enumerate(plan.groups):
            n, table, K, G, maxp, L, ext = grp
            if kept_mode and keeps_summaries(grp, X.dtype):
                ws[gi] = ops.lpc_ragged_keep_(X, AA, Z, table, y, K, G, maxp, n, L, ext)
            else:
                ops.lpc_ragged_(X, AA, Z, table, y, K, G, maxp, n, L, ext)
        outs = []
        for _ in range(steps):
            dx, da, dzi = torch.empty_like(X), torch.empty_like(AA), torch.empty_like(Z)
            for gi, grp in enumerate(plan.groups):
                n, table, K, G, maxp, L, ext = grp
                ops.lpc_ragged_backward_(Gd, AA, Z, y, table, dx, da, dzi, K, G, maxp, n, L, True, True, True,
                                         ext, ws.get(gi))
            outs.append((dx, da, dzi))
        return y, outs

    y0, [(dx0, da0, dz0)] = run(False)
    y1, [(dx1, da1, dz1), (dx2, da2, dz2)] = run(True, steps=2)
    assert torch.equal(y0, y1)
    for u, v in ((dx0, dx1), (da0, da1), (dz0, dz1)):
        assert rel(host(u), host(v)) <= 1e-5
    for u, v in ((dx1, dx2), (da1, da2), (dz1, dz2)):
        assert torch.equal(u, v)

    # graph replay of the kept step
    def step():
        return run(True)[1][0]

    eager = [t.clone() for t in step()]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = step()
    for r in range(3):
        torch.randn(1 << 20, device="cuda").mul_(3.0)
        graph.replay()
        torch.cuda.synchronize()
        for e, o in zip(eager, outs):
            assert torch.equal(e, o), r


def test_ragged_backward_partial_gradients():
    """dx only, dA only, dzi only and dx with dzi: every requested output equals
    the full backward's bit for bit, at orders 2 to 32 (the adjoint variants
    without dx or dA stores, in the quad and group kernels and the staged
    writes of orders 16 and 32), with chunks ending past a signal's end and rows
    that are not 16-byte aligned. The full backward is checked against FP64."""
    from torchlpc_amd import RaggedPlan

    rng = np.random.default_rng(31)
    lengths = [7, 700, 3001, 64, 1025, 5000, 333, 2048, 129]
    orders = [1, 2, 16, 32, 8, 32, 4, 16, 8]
    sigs = _mixed(rng, lengths, orders)
    x, a, zi, g = _pack(sigs, np.float32)
    X, AA, Z, Gd = dev(x), dev(a), dev(zi), dev(g)
    plan = RaggedPlan(lengths, orders, X.device, X.dtype, chunk_len=128)
    ops = torch.ops.torchlpc_amd
    y = torch.empty_like(X)
    for n, table, K, G, maxp, L, ext in plan.groups:
        ops.lpc_ragged_(X, AA, Z, table, y, K, G, maxp, n, L, ext)

    def backward(need_x, need_a, need_zi):
        outs = [torch.full_like(t, float("nan")) for t in (X, AA, Z)]
        for n, table, K, G, maxp, L, ext in plan.groups:
            ops.lpc_ragged_backward_(Gd, AA, Z, y, table, *outs, K, G, maxp, n, L, need_x, need_a, need_zi, ext,
                                     None)
        return outs

    full = backward(True, True, True)
    _check(sigs, np.float32, plan, y, *full, REL_TOL[np.float32])
    for flags in ((True, False, False), (False, True, False), (False, False, True), (True, False, True)):
        part = backward(*flags)
        for want, f, p in zip(flags, full, part):
            if want:
                assert torch.equal(f, p), flags
