"""Per-signal time-invariant kernels (orders 16..32, float32, strategy 2): block
products on MFMA, float64 state transitions. y, dx, da and dzi against the FP64
oracle across block (64), superblock (512), column-tile (2048) and segment (8192)
boundaries, the requested-gradient subsets, and HIP-graph replay."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.gpu_util import dev, host, rel, up

pytestmark = pytest.mark.gpu


def ops():
    import torchlpc_amd  # noqa: F401

    return torch.ops.torchlpc_amd


def _inputs(rng, B, T, N, kmax=None):
    kmax = min(0.8, 6.0 / N) if kmax is None else kmax
    a = O.step_up(rng.uniform(-kmax, kmax, (B, N))).astype(np.float32)
    x = rng.standard_normal((B, T)).astype(np.float32)
    zi = rng.standard_normal((B, N)).astype(np.float32)
    g = rng.standard_normal((B, T)).astype(np.float32)
    return x, a, zi, g


def _oracle(x, a, zi, g):
    B, T = x.shape
    N = a.shape[1]
    A3 = np.ascontiguousarray(np.broadcast_to(up(a, np.float32)[:, None, :], (B, T, N)))
    y64 = O.lpc(up(x, np.float32), A3, up(zi, np.float32))
    rx, rA, rzi = O.lpc_backward(up(g, np.float32), A3, up(zi, np.float32), y64)
    return y64, rx, rA.sum(1), rzi


@pytest.mark.parametrize("B,T,N", [(3, 1, 16), (2, 5, 20), (3, 63, 17), (2, 64, 32), (3, 65, 32), (2, 511, 32),
                                   (3, 2049, 32), (2, 8192, 24), (2, 8193, 32), (3, 20000, 32), (2, 16384, 16)])
def test_ti_mfma_forward_backward(B, T, N):
    rng = np.random.default_rng(B * 7919 + T * 31 + N)
    x, a, zi, g = _inputs(rng, B, T, N)
    y = ops().lpc(dev(x), dev(a), dev(zi), 2, 0)
    dx, da, dzi = ops().lpc_backward(dev(g), dev(a), dev(zi), y, True, True, True, 2, 0)
    y64, rx, rda, rzi = _oracle(x, a, zi, g)
    errs = [rel(host(y), y64), rel(host(dx), rx), rel(host(da), rda), rel(host(dzi), rzi)]
    print((B, T, N), errs, flush=True)
    assert max(errs) <= 1e-5, errs


def test_ti_mfma_poles_near_the_unit_circle():
    """k ~ U(-0.8, 0.8) at order 32: poles within about 1e-6 of the unit circle,
    where float32 state transitions would lose 1e-4; the float64 transitions keep
    every output within 1e-5."""
    rng = np.random.default_rng(1005)
    x, a, zi, g = _inputs(rng, 8, 16384, 32, kmax=0.8)
    zi *= 0.1
    y = ops().lpc(dev(x), dev(a), dev(zi), 2, 0)
    dx, da, dzi = ops().lpc_backward(dev(g), dev(a), dev(zi), y, True, True, True, 2, 0)
    y64, rx, rda, rzi = _oracle(x, a, zi, g)
    errs = [rel(host(y), y64), rel(host(dx), rx), rel(host(da), rda), rel(host(dzi), rzi)]
    print("near-unit-circle", errs, flush=True)
    assert max(errs) <= 1e-5, errs


@pytest.mark.parametrize("need", [(True, False, False), (False, True, False), (False, False, True), (True, False, True)])
def test_ti_mfma_need_flags(need):
    """Only the requested gradients are written; without da the kernel never reads
    y (passed as NaNs here)."""
    rng = np.random.default_rng(11)
    B, T, N = 3, 5000, 32
    x, a, zi, g = _inputs(rng, B, T, N)
    y64, rx, rda, rzi = _oracle(x, a, zi, g)
    y = dev(y64.astype(np.float32)) if need[1] else torch.full((B, T), float("nan"), device="cuda")
    outs = ops().lpc_backward(dev(g), dev(a), dev(zi), y, *need, 2, 0)
    for o, r, n in zip(outs, (rx, rda, rzi), need):
        if n:
            assert rel(host(o), r) <= 1e-5
        else:
            assert o.numel() == 0


def test_ti_mfma_matches_chunked_path():
    """The MFMA kernels and the chunked time-invariant scan agree within the bar
    (different association, same filter)."""
    rng = np.random.default_rng(3)
    x, a, zi, g = (dev(v) for v in _inputs(rng, 4, 12345, 28))
    y2 = ops().lpc(x, a, zi, 2, 0)
    y1 = ops().lpc(x, a, zi, 1, 0)
    assert float((y2 - y1).norm() / y1.norm()) <= 1e-5
    g2 = ops().lpc_backward(g, a, zi, y2, True, True, True, 2, 0)
    g1 = ops().lpc_backward(g, a, zi, y1, True, True, True, 1, 0)
    for u, v in zip(g2, g1):
        assert float((u - v).norm() / v.norm()) <= 1e-5


def test_hip_graph_capture_replays_ti_mfma():
    """Forward and backward of the MFMA kernels captured in one HIP graph: every
    replay equals the eager result bit for bit."""
    rng = np.random.default_rng(32)
    X, AA, Z, G = (dev(v) for v in _inputs(rng, 5, 9000, 32))

    def step():
        y = ops().lpc(X, AA, Z, 2, 0)
        return (y, *ops().lpc_backward(G, AA, Z, y, True, True, True, 2, 0))

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
