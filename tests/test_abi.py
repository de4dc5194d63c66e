"""The C-ABI library loads, exports every declared entry point and validates
arguments (CPU only: no call here reaches a kernel launch)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest
import torch  # noqa: F401  (maps torch's HIP runtime before the library)

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "torchlpc_amd.h"
LIB = ROOT / "torchlpc_amd" / "libtorchlpc_amd_hip.so"

TLPC_F32, TLPC_F64, TLPC_C64, TLPC_C128 = 0, 1, 2, 3
TLPC_EINVAL, TLPC_EDTYPE, TLPC_EWORKSPACE = -1, -2, -3
NEED_DX, NEED_DA, NEED_DZI, TI, DA_FULL = 1, 2, 4, 8, 16


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tlpc_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        pytest.fail(f"{LIB} is not built (python -m torchlpc_amd._build)")
    L = ctypes.CDLL(str(LIB))
    i64, i32, p, sz = ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t
    L.tlpc_error_string.restype = ctypes.c_char_p
    L.tlpc_lpc_forward_workspace.restype = sz
    L.tlpc_lpc_forward_workspace.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    L.tlpc_lpc_backward_workspace.restype = sz
    L.tlpc_lpc_backward_workspace.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    L.tlpc_scan_workspace.restype = sz
    L.tlpc_scan_workspace.argtypes = [ctypes.c_int, i64, i64, i32, i32]
    L.tlpc_lpc_forward.argtypes = [ctypes.c_int, p, p, p, p, i64, i64, i32, i32, i32, i32, p, sz, p]
    L.tlpc_lpc_backward.argtypes = [ctypes.c_int, p, p, p, p, p, p, p, i64, i64, i32, i32, i32, i32, p, sz, p]
    L.tlpc_scan.argtypes = [ctypes.c_int, p, p, p, p, i64, i64, i32, p, sz, p]
    return L


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 9, names
    for name in names:
        assert hasattr(lib, name), f"{name} declared in {HEADER.name} but not exported"


def test_error_strings(lib):
    assert lib.tlpc_error_string(0) == b"ok"
    assert lib.tlpc_error_string(TLPC_EWORKSPACE) == b"workspace too small"
    assert lib.tlpc_abi_version() == 1


def test_workspace_queries(lib):
    # chunked: B*(P-1)*(N*N+N) summaries + B*P*N carries (256-byte aligned), kept in
    # double precision for float data
    B, T, N = 256, 48000, 16
    ws = lib.tlpc_lpc_forward_workspace(TLPC_F32, B, T, N, 0, 1, 512)
    P = -(-T // 512)
    summ = B * (P - 1) * (N * N + N) * 8
    carry = B * P * N * 8
    assert ws == -(-summ // 256) * 256 + -(-carry // 256) * 256
    assert lib.tlpc_lpc_forward_workspace(TLPC_F64, B, T, N, 0, 1, 512) == ws
    assert lib.tlpc_lpc_forward_workspace(TLPC_C64, B, T, N, 0, 1, 512) == 2 * ws
    # time-invariant reduced da needs per-chunk partials; the full per-sample dA does not
    wb = lib.tlpc_lpc_backward_workspace(TLPC_F32, 8, 4096, 4, NEED_DA | TI, 1, 256)
    wf = lib.tlpc_lpc_backward_workspace(TLPC_F32, 8, 4096, 4, NEED_DA | TI | DA_FULL, 1, 256)
    assert wb > wf
    # invalid queries report 0
    assert lib.tlpc_lpc_forward_workspace(7, B, T, N, 0, 0, 0) == 0
    assert lib.tlpc_lpc_forward_workspace(TLPC_F32, B, T, 0, 0, 0, 0) == 0
    assert lib.tlpc_scan_workspace(TLPC_F32, 4, 1 << 20, 0, 0) > 0
    # chains of more than 1536 links at orders <= 4 add segment maps and states:
    # B * nseg * (N*N + 2N) wide elements, nseg = ceil(links / 512)
    B, T, N, L = 4, 16 * 4100, 2, 16
    P = T // L
    nseg = -(-(P - 1) // 512)
    summ = B * (P - 1) * (N * N + N) * 8
    carry = B * P * N * 8
    segs = B * nseg * (N * N + 2 * N) * 8
    up = lambda v: -(-v // 256) * 256
    assert lib.tlpc_lpc_forward_workspace(TLPC_F32, B, T, N, 0, 1, L) == up(summ) + up(carry) + up(segs)
    # shorter chains (1536 links) keep one workgroup per chain and no segment space
    T = 16 * 1537
    P = T // L
    assert lib.tlpc_lpc_forward_workspace(TLPC_F32, B, T, N, 0, 1, L) == up(B * (P - 1) * 6 * 8) + up(B * P * 2 * 8)


def test_summary_bytes(lib):
    """The forward/backward shared region: the float64 transition matrices and
    zero-state responses of chunks 0..P-2, for float32 order-16 time-varying
    filters only."""
    i64, i32 = ctypes.c_int64, ctypes.c_int32
    f = lib.tlpc_lpc_summary_bytes
    f.restype = ctypes.c_size_t
    f.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    B, T, N, L = 256, 48000, 16, 384
    P = -(-T // L)
    summ = B * (P - 1) * (N * N + N) * 8
    assert f(TLPC_F32, B, T, N, 0, 1, L) == -(-summ // 256) * 256
    fwd = lib.tlpc_lpc_forward_workspace(TLPC_F32, B, T, N, 0, 1, L)
    bwd = lib.tlpc_lpc_backward_workspace(TLPC_F32, B, T, N, NEED_DX | NEED_DA | NEED_DZI, 1, L)
    assert fwd < bwd  # the backward adds its zero-state responses after the carries
    assert f(TLPC_F64, B, T, N, 0, 1, L) == 0
    assert f(TLPC_F32, B, T, 8, 0, 1, L) == 0
    assert f(TLPC_F32, B, T, N, TI, 1, L) == 0
    assert f(TLPC_F32, B, T, N, 0, 3, L) == 0  # serial strategy keeps no summaries


def test_argument_validation_without_launch(lib):
    buf = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    assert lib.tlpc_lpc_forward(9, buf, buf, buf, buf, 2, 8, 2, 0, 0, 0, None, 0, None) == TLPC_EDTYPE
    assert lib.tlpc_lpc_forward(TLPC_F32, buf, buf, buf, buf, 2, 8, 0, 0, 0, 0, None, 0, None) == TLPC_EINVAL
    assert lib.tlpc_lpc_forward(TLPC_F32, buf, None, buf, buf, 2, 8, 2, 0, 0, 0, None, 0, None) == TLPC_EINVAL
    # the chunked forward needs a workspace
    assert lib.tlpc_lpc_forward(TLPC_F32, buf, buf, buf, buf, 2, 4096, 2, 0, 1, 64, None, 0, None) == TLPC_EWORKSPACE
    # a requested gradient needs its output pointer
    assert (
        lib.tlpc_lpc_backward(TLPC_F32, buf, buf, buf, buf, None, None, None, 2, 8, 2, NEED_DX, 0, 0, None, 0, None)
        == TLPC_EINVAL
    )
    # B == 0 is a no-op
    assert lib.tlpc_lpc_forward(TLPC_F32, None, buf, None, None, 0, 8, 2, 0, 0, 0, None, 0, None) == 0
    # an empty signal (T == 0): per-sample buffers have no storage, so null is valid
    # there; a time-invariant A is per signal and still required
    assert lib.tlpc_lpc_forward(TLPC_F32, None, None, buf, None, 2, 0, 4, 0, 0, 0, None, 0, None) == 0
    assert lib.tlpc_lpc_forward(TLPC_F32, None, None, buf, None, 2, 0, 4, TI, 0, 0, None, 0, None) == TLPC_EINVAL
    assert lib.tlpc_scan(TLPC_F32, None, None, buf, None, 2, 0, 0, None, 0, None) == 0


def test_recompute_queries_and_validation(lib):
    i64, i32, p, sz = ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t
    lib.tlpc_lpc_recompute_supported.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    f = lib.tlpc_lpc_backward_recompute
    f.argtypes = [ctypes.c_int, p, p, p, p, p, sz, p, p, p, i64, i64, i32, i32, i32, i32, p, sz, p]
    assert lib.tlpc_lpc_recompute_supported(TLPC_F32, 4, 1 << 20, 2, NEED_DX | NEED_DA, 0, 0) == 1
    assert lib.tlpc_lpc_recompute_supported(TLPC_F64, 4, 1 << 20, 4, NEED_DX, 1, 128) == 1
    assert lib.tlpc_lpc_recompute_supported(TLPC_F32, 4, 1 << 20, 5, NEED_DX, 0, 0) == 0
    assert lib.tlpc_lpc_recompute_supported(TLPC_F32, 4, 1 << 20, 2, NEED_DX | TI, 0, 0) == 0
    assert lib.tlpc_lpc_recompute_supported(TLPC_C64, 4, 1 << 20, 2, NEED_DX, 0, 0) == 0
    assert lib.tlpc_lpc_recompute_supported(TLPC_F32, 4, 1 << 20, 2, NEED_DX, 1, 512) == 0
    buf = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation first
    # order 16 is not served
    assert f(TLPC_F32, buf, buf, buf, buf, buf, 1 << 30, buf, None, None, 2, 4096, 16, NEED_DX, 0, 0, buf, 1 << 30, None) == TLPC_EINVAL
    # no forward input
    assert f(TLPC_F32, buf, buf, buf, None, buf, 1 << 30, buf, None, None, 2, 4096, 2, NEED_DX, 0, 0, buf, 1 << 30, None) == TLPC_EINVAL
    # forward workspace too small
    assert f(TLPC_F32, buf, buf, buf, buf, buf, 8, buf, None, None, 2, 4096, 2, NEED_DX, 0, 0, buf, 1 << 30, None) == TLPC_EWORKSPACE
