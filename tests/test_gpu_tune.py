"""The autotuner on the GPU: every candidate it times gives results within the bar,
the stored choice is what the autograd Function then runs, and lookups are keyed
on the exact shape."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.gpu_util import dev, host, rel, stable_inputs, up

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("B,T,N", [(4, 6000, 16), (3, 5000, 4)])
def test_tune_choice_is_used_and_every_candidate_is_accurate(B, T, N, tmp_path, monkeypatch):
    import torchlpc_amd
    from torchlpc_amd import autotune

    monkeypatch.setenv("TORCHLPC_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.clear()
    ops = torch.ops.torchlpc_amd
    rng = np.random.default_rng(B * T + N)
    x, A, zi = stable_inputs(rng, B, T, N, np.float32, hop=240)
    g = rng.standard_normal((B, T)).astype(np.float32)
    X, AA, Z, G = dev(x), dev(A), dev(zi), dev(g)
    res = autotune.tune(X, AA, Z, grad=G, reps=2, persist=True)
    choice = tuple(res["choice"])
    assert autotune.lookup("lpc_forward", X, AA) == choice
    assert autotune.lookup("lpc_backward", G, AA) == choice
    assert (tmp_path / "tune.json").exists()
    assert len(res["ms"]) >= 3 and res["heuristic_ms"] > 0

    y64 = O.lpc(up(x, np.float32), up(A, np.float32), up(zi, np.float32))
    ref = O.lpc_backward(up(g, np.float32), up(A, np.float32), up(zi, np.float32), y64)
    for key in res["ms"]:
        s, L = (int(v) for v in key.split(","))
        y, (dx, dA, dzi) = autotune.step_fn(X, AA, Z, G, (s, L))()
        errs = [rel(host(y), y64)] + [rel(host(o), r) for o, r in zip((dx, dA, dzi), ref)]
        assert max(errs) <= 1e-5, (key, errs)

    # the autograd Function runs the stored choice: bitwise equal to the explicit call
    Xg, Ag, Zg = (t.clone().requires_grad_() for t in (X, AA, Z))
    yy = torchlpc_amd.LPC.apply(Xg, Ag, Zg)
    yy.backward(G)
    y_ref, ws = ops.lpc_keep(X, AA, Z, *choice)
    grads = ops.lpc_backward(G, AA, Z, y_ref, True, True, True, *choice, ws if ws.numel() else None)
    assert torch.equal(yy.detach(), y_ref)
    for got, want in zip((Xg.grad, Ag.grad, Zg.grad), grads):
        assert torch.equal(got, want)

    # a different length has no stored choice
    assert autotune.lookup("lpc_forward", X[:, : T - 16], AA[:, : T - 16]) == (0, 0)
    autotune.clear()
