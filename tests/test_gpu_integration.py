"""Integration with PyTorch's operator tooling and with callers of the C ABI.

* torch.library.opcheck on every op of the namespace (schema, fake kernels,
  autograd registration, AOT dispatch);
* torch.compile of a function that filters and back-propagates, against eager;
* tlpc_lpc_forward / tlpc_lpc_backward launched through ctypes exactly as the
  integration guide shows, against the torch ops (bit for bit: same plan).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.gpu_util import dev, host, rel, stable_inputs, up

pytestmark = pytest.mark.gpu


def ops():
    import torchlpc_amd  # noqa: F401

    return torch.ops.torchlpc_amd


@pytest.mark.parametrize("N,T", [(16, 700), (4, 500)])
def test_opcheck_lpc_ops(N, T):
    rng = np.random.default_rng(N)
    x, A, zi = (dev(v) for v in stable_inputs(rng, 3, T, N, np.float64, hop=240))
    g = dev(rng.standard_normal((3, T)))
    y = ops().lpc(x, A, zi)
    # the ops are plain kernels (no autograd kernel: differentiation is the LPC
    # Function's), so they are checked on tensors that do not require grad
    torch.library.opcheck(ops().lpc.default, (x, A, zi, 0, 0))
    torch.library.opcheck(ops().lpc_keep.default, (x, A, zi, 0, 0))
    torch.library.opcheck(ops().lpc_backward.default, (g, A, zi, y, True, True, True, 0, 0, None))


def test_opcheck_scan_ops():
    rng = np.random.default_rng(5)
    B, T = 3, 600
    x = dev(rng.standard_normal((B, T)))
    w = dev(rng.uniform(-0.9, 0.9, (B, T)))
    init = dev(rng.standard_normal(B))
    h = ops().scan(x, w, init)
    g = dev(rng.standard_normal((B, T)))
    torch.library.opcheck(ops().scan.default, (x, w, init, 0))
    torch.library.opcheck(ops().scan_backward.default, (g, w, init, h, True, True, True, 0))


def test_torch_compile_forward_and_backward():
    """sample_wise_lpc inside a compiled function: outputs and gradients equal
    the eager ones. Dynamo does not trace autograd Functions that define a jvp,
    so LPC runs between compiled graphs, on the same kernels."""
    import torchlpc_amd

    rng = np.random.default_rng(11)
    B, T, N = 3, 1500, 16
    x, A, zi = stable_inputs(rng, B, T, N, np.float32, kmax=0.8, hop=240)
    w = dev(rng.standard_normal((B, T)).astype(np.float32))

    def loss(x, A, zi):
        y, zf = torchlpc_amd.sample_wise_lpc(x, A, zi, return_zf=True)
        return (y * w).sum() + zf.sum()

    results = []
    for fn in (loss, torch.compile(loss, dynamic=False)):
        X, AA, Z = (dev(v).requires_grad_() for v in (x, A, zi))
        val = fn(X, AA, Z)
        val.backward()
        results.append((val.detach(), X.grad, AA.grad, Z.grad))
    for e, c in zip(*results):
        assert rel(host(c), host(e)) <= 1e-6


def _lib():
    import torchlpc_amd
    from torchlpc_amd import _ext

    assert torchlpc_amd.EXTENSION_LOADED, torchlpc_amd.LOAD_ERROR
    return ctypes.CDLL(str(_ext.HIP_LIB))


@pytest.mark.parametrize("B,T,N", [(4, 3000, 16), (3, 2000, 5), (256, 2048, 16)])
def test_c_abi_launch_through_ctypes(B, T, N):
    """The binding of the integration guide: workspace query, forward and fused
    backward on torch-owned device buffers and the current stream."""
    lib = _lib()
    i64, i32, p, sz = ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t
    for name in ("tlpc_lpc_forward_workspace", "tlpc_lpc_backward_workspace"):
        getattr(lib, name).restype = sz
        getattr(lib, name).argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    lib.tlpc_lpc_forward.argtypes = [ctypes.c_int, p, p, p, p, i64, i64, i32, i32, i32, i32, p, sz, p]
    lib.tlpc_lpc_backward.argtypes = [ctypes.c_int, p, p, p, p, p, p, p, i64, i64, i32, i32, i32, i32, p, sz, p]
    lib.tlpc_error_string.restype = ctypes.c_char_p

    rng = np.random.default_rng(B + T + N)
    xh, Ah, zih = stable_inputs(rng, B, T, N, np.float32, hop=240)
    gh = rng.standard_normal((B, T)).astype(np.float32)
    x, A, zi, g = dev(xh), dev(Ah), dev(zih), dev(gh)  # contiguous, as the ABI requires
    stream = torch.cuda.current_stream().cuda_stream

    nbytes = lib.tlpc_lpc_forward_workspace(0, B, T, N, 0, 0, 0)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=x.device)
    y = torch.empty_like(x)
    rc = lib.tlpc_lpc_forward(0, x.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(),
                              B, T, N, 0, 0, 0, ws.data_ptr(), nbytes, stream)
    assert rc == 0, lib.tlpc_error_string(rc)

    flags = 1 | 2 | 4  # TLPC_NEED_DX | TLPC_NEED_DA | TLPC_NEED_DZI
    nb = lib.tlpc_lpc_backward_workspace(0, B, T, N, flags, 0, 0)
    wsb = torch.empty(max(nb, 1), dtype=torch.uint8, device=x.device)
    dx, dA, dzi = torch.empty_like(x), torch.empty_like(A), torch.empty_like(zi)
    rc = lib.tlpc_lpc_backward(0, g.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(),
                               dx.data_ptr(), dA.data_ptr(), dzi.data_ptr(), B, T, N, flags, 0, 0,
                               wsb.data_ptr(), nb, stream)
    assert rc == 0, lib.tlpc_error_string(rc)
    torch.cuda.synchronize()

    assert torch.equal(y, ops().lpc(x, A, zi))
    for got, want in zip((dx, dA, dzi), ops().lpc_backward(g, A, zi, y, True, True, True)):
        assert torch.equal(got, want)
    y64 = O.lpc(up(xh, np.float32), up(Ah, np.float32), up(zih, np.float32))
    ref = O.lpc_backward(up(gh, np.float32), up(Ah, np.float32), up(zih, np.float32), y64)
    assert rel(host(y), y64) <= 1e-5
    for got, r in zip((dx, dA, dzi), ref):
        assert rel(host(got), r) <= 1e-5
