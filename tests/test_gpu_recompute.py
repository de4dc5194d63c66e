"""The adjoint that rebuilds y from the forward's chunk carries (long signals,
orders up to 4): equal to the y-reading adjoint given the forward's y, within
the bars of the FP64 oracle at the full B=4, T=2^20, N=2 size in float32 and
float64, selected by the autograd Function for long signals, and HIP-graph
replayable."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.gpu_util import REL_TOL, dev, host, rel, stable_inputs, up

pytestmark = pytest.mark.gpu


def ops():
    import torchlpc_amd  # noqa: F401

    return torch.ops.torchlpc_amd


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("B,T,N,chunk", [(3, 5000, 2, 0), (2, 4099, 1, 0), (3, 3001, 3, 64), (2, 20000, 4, 256),
                                         (4, 777, 2, 16), (1, 70000, 2, 128)])
def test_recompute_equals_saved_y_backward(dtype, B, T, N, chunk):
    rng = np.random.default_rng(B * 1009 + T + N)
    x, A, zi = stable_inputs(rng, B, T, N, dtype, hop=240)
    g = rng.standard_normal((B, T)).astype(dtype)
    X, AA, Z, G = dev(x), dev(A), dev(zi), dev(g)
    y, ws = ops().lpc_carries(X, AA, Z, 1, chunk)
    assert torch.equal(y, ops().lpc(X, AA, Z, 1, chunk))
    rc = ops().lpc_backward_recompute(G, AA, Z, X, ws, True, True, True, 1, chunk)
    ref = ops().lpc_backward(G, AA, Z, y, True, True, True, 1, chunk)
    for u, v in zip(rc, ref):
        assert rel(host(u), host(v)) <= (1e-7 if dtype == np.float32 else 1e-14)
    y64 = O.lpc(up(x, dtype), up(A, dtype), up(zi, dtype))
    r64 = O.lpc_backward(up(g, dtype), up(A, dtype), up(zi, dtype), y64)
    for u, r in zip(rc, r64):
        assert rel(host(u), r) <= REL_TOL[dtype]


@pytest.mark.parametrize("need", [(True, False, False), (False, True, False), (False, False, True)])
def test_recompute_need_flags(need):
    rng = np.random.default_rng(7)
    x, A, zi = stable_inputs(rng, 2, 6000, 2, np.float32, hop=240)
    g = rng.standard_normal((2, 6000)).astype(np.float32)
    X, AA, Z, G = dev(x), dev(A), dev(zi), dev(g)
    y, ws = ops().lpc_carries(X, AA, Z, 1, 0)
    rc = ops().lpc_backward_recompute(G, AA, Z, X, ws, *need, 1, 0)
    ref = ops().lpc_backward(G, AA, Z, y, *need, 1, 0)
    for u, v, n in zip(rc, ref, need):
        if n:
            assert rel(host(u), host(v)) <= 1e-7
        else:
            assert u.numel() == 0


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_drifting_b4_t1m_n2_autograd_rebuilds_y(dtype, tol):
    """The full-size long-signal workload through LPC.apply: the Function keeps
    the forward's carries (not y) and the gradients match the FP64 oracle."""
    import torchlpc_amd

    x, A, zi, w = (v.astype(dtype) for v in O.make_drifting_b4_t1m_n2())
    X, AA, Z = (dev(v).requires_grad_() for v in (x, A, zi))
    saved = []
    with torch.autograd.graph.saved_tensors_hooks(lambda t: saved.append(t) or t, lambda t: t):
        y = torchlpc_amd.LPC.apply(X, AA, Z)
    assert not any(t.data_ptr() == y.data_ptr() for t in saved), "y was saved"
    assert any(t.dtype == torch.uint8 for t in saved), "no carries kept"
    kept = sum(t.numel() for t in saved if t.dtype == torch.uint8)
    print("kept carries", kept, "bytes instead of y", y.numel() * y.element_size(), flush=True)
    assert kept < y.numel() * y.element_size() / 4
    y.backward(dev(w))
    up_ = lambda v: v.astype(np.float64)  # noqa: E731
    y64 = O.lpc(up_(x), up_(A), up_(zi))
    rx, rA, rzi = O.lpc_backward(up_(w), up_(A), up_(zi), y64)
    errs = [rel(host(y.detach()), y64), rel(host(X.grad), rx), rel(host(AA.grad), rA), rel(host(Z.grad), rzi)]
    print("b4-t1m-n2 rebuild-y errors", errs, flush=True)
    assert max(errs) <= tol, errs


def test_hip_graph_capture_replays_recompute():
    rng = np.random.default_rng(19)
    x, A, zi = stable_inputs(rng, 3, 9000, 2, np.float32, hop=240)
    g = rng.standard_normal((3, 9000)).astype(np.float32)
    X, AA, Z, G = dev(x), dev(A), dev(zi), dev(g)

    def step():
        y, ws = ops().lpc_carries(X, AA, Z, 1, 0)
        return (y, *ops().lpc_backward_recompute(G, AA, Z, X, ws, True, True, True, 1, 0))

    eager = [t.clone() for t in step()]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = step()
    for r in range(4):
        junk = torch.randn(1 << 20, device="cuda") * 2.0
        X.mul_(1.0)
        graph.replay()
        torch.cuda.synchronize()
        for e, o in zip(eager, outs):
            assert torch.equal(e, o), (r, float((e - o).abs().max()))
        del junk
