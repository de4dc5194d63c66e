"""The allocation-free forward entry point, tlpc_lpc_forward_noalloc, through the C ABI.

It must give bit for bit what tlpc_lpc_forward gives (the same plan and kernels,
with the workspace in the library's static device scratch), on two calls in a
row where the second is smaller, and reject sizes beyond its scratch before
launching anything.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from tests.gpu_util import dev, stable_inputs

pytestmark = pytest.mark.gpu

TLPC_F32, TLPC_F64 = 0, 1
TLPC_ESCRATCH = -6
SCRATCH_BYTES = 64 << 20


def _lib():
    import torchlpc_amd
    from torchlpc_amd import _ext

    assert torchlpc_amd.EXTENSION_LOADED, torchlpc_amd.LOAD_ERROR
    lib = ctypes.CDLL(str(_ext.HIP_LIB))
    i64, i32, p = ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p
    lib.tlpc_lpc_forward_noalloc.restype = ctypes.c_int
    lib.tlpc_lpc_forward_noalloc.argtypes = [ctypes.c_int, p, p, p, p, i64, i64, i32, i32, i32, i32, p]
    lib.tlpc_lpc_forward_workspace.restype = ctypes.c_size_t
    lib.tlpc_lpc_forward_workspace.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    lib.tlpc_error_string.restype = ctypes.c_char_p
    return lib


def _noalloc(lib, x, A, zi, flags=0):
    y = torch.full_like(x, float("nan"))
    code = TLPC_F32 if x.dtype == torch.float32 else TLPC_F64
    B, T = x.shape
    N = A.shape[-1]
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.tlpc_lpc_forward_noalloc(
        code, x.data_ptr(), A.data_ptr(), zi.data_ptr(), y.data_ptr(), B, T, N, flags, 0, 0, stream
    )
    assert rc == 0, lib.tlpc_error_string(rc)
    return y


@pytest.mark.parametrize(
    "shapes",
    [
        # order 16 float32 (quad kernels), then a smaller order-4 float64 call
        ((64, 20000, 16, np.float32), (8, 3000, 4, np.float64)),
        # generic order-8 float32, then a smaller order-8 float32 call (the scratch keeps
        # the first call's contents)
        ((32, 9000, 8, np.float32), (3, 777, 8, np.float32)),
    ],
)
def test_noalloc_forward_matches_forward(shapes):
    import torchlpc_amd  # noqa: F401

    lib = _lib()
    ops = torch.ops.torchlpc_amd
    rng = np.random.default_rng(7)
    for B, T, N, dt in shapes:
        print(f"noalloc forward B={B} T={T} N={N} {np.dtype(dt).name}", flush=True)
        # the C ABI takes contiguous buffers (dev() makes them so)
        x, A, zi = (dev(v) for v in stable_inputs(rng, B, T, N, dt, hop=240))
        code = TLPC_F32 if dt == np.float32 else TLPC_F64
        assert 0 < lib.tlpc_lpc_forward_workspace(code, B, T, N, 0, 0, 0) <= SCRATCH_BYTES
        y_ref = ops.lpc(x, A, zi)
        y = _noalloc(lib, x, A, zi)
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref)


def test_noalloc_forward_time_invariant():
    lib = _lib()
    ops = torch.ops.torchlpc_amd
    rng = np.random.default_rng(8)
    B, T, N = 16, 5000, 32
    x = dev(rng.standard_normal((B, T)).astype(np.float32))
    a = dev(rng.uniform(-0.02, 0.02, (B, N)).astype(np.float32))
    zi = dev(rng.standard_normal((B, N)).astype(np.float32))
    y = _noalloc(lib, x, a, zi, flags=8)
    torch.cuda.synchronize()
    assert torch.equal(y, ops.lpc(x, a, zi))


def test_noalloc_forward_rejects_oversize():
    """Sizes whose workspace exceeds the static scratch return TLPC_ESCRATCH (here
    the chunked strategy, whose summaries grow with the batch); the check needs no
    device memory (the pointers are never read)."""
    lib = _lib()
    B, T, N, chunked = 4096, 48000, 16, 1
    assert lib.tlpc_lpc_forward_workspace(TLPC_F32, B, T, N, 0, chunked, 0) > SCRATCH_BYTES
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.tlpc_lpc_forward_noalloc(TLPC_F32, 1, 1, 1, 1, B, T, N, 0, chunked, 0, stream)
    assert rc == TLPC_ESCRATCH
