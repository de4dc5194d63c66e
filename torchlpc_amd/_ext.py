"""Loader for the native operator library (``_C.so`` + ``libtorchlpc_amd_hip.so``).

Reference counterpart: ``torchlpc/__init__.py:17-23`` (``from . import _C``).
Unlike the reference there is no fallback: when the library is missing, every
call into the filter raises ``RuntimeError`` naming the build command.
"""
from __future__ import annotations

from pathlib import Path

import torch

PKG = Path(__file__).resolve().parent
OPS_LIB = PKG / "_C.so"
HIP_LIB = PKG / "libtorchlpc_amd_hip.so"

_OPS = (
    "lpc",
    "lpc_keep",
    "lpc_backward",
    "lpc_carries",
    "lpc_backward_recompute",
    "lpc_ragged_plan",
    "lpc_ragged_",
    "lpc_ragged_keep_",
    "lpc_ragged_backward_",
    "scan",
    "scan_backward",
)

EXTENSION_LOADED = False
LOAD_ERROR: str | None = None


def _load() -> None:
    global EXTENSION_LOADED, LOAD_ERROR
    if EXTENSION_LOADED:
        return
    if not OPS_LIB.exists():
        LOAD_ERROR = f"{OPS_LIB} not found; build it with `python -m torchlpc_amd._build`"
        return
    try:
        torch.ops.load_library(str(OPS_LIB))
    except (OSError, RuntimeError) as e:  # pragma: no cover - environment specific
        LOAD_ERROR = f"failed to load {OPS_LIB}: {e}"
        return
    missing = [op for op in _OPS if not hasattr(torch.ops.torchlpc_amd, op)]
    if missing:
        LOAD_ERROR = (
            f"{OPS_LIB} predates this package (missing ops {missing}); rebuild it with "
            "`python -m torchlpc_amd._build`"
        )
        return
    EXTENSION_LOADED = True
    _register_fake()


_DTYPE_CODE = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}
_HIPLIB = None


def _hiplib():
    """The C-ABI library through ctypes, for its host-side queries (opened once)."""
    global _HIPLIB
    if _HIPLIB is None:
        import ctypes

        _HIPLIB = ctypes.CDLL(str(HIP_LIB))
    return _HIPLIB


def keep_workspace_bytes(x: torch.Tensor, A: torch.Tensor, strategy: int, chunk_len: int) -> int:
    """Bytes of the workspace ``lpc_keep`` returns for these inputs (0: nothing kept).

    Host-side size queries of the C ABI (``tlpc_lpc_summary_bytes`` and the
    forward/backward workspace queries); they launch nothing.
    """
    import ctypes

    lib = _hiplib()
    i64, i32 = ctypes.c_int64, ctypes.c_int32
    for name in ("tlpc_lpc_summary_bytes", "tlpc_lpc_forward_workspace", "tlpc_lpc_backward_workspace"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_size_t
        fn.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    dt = _DTYPE_CODE.get(x.dtype)
    if dt is None:
        return 0
    B, T = x.shape
    N = A.shape[-1]
    ti = A.dim() == 2 or (A.shape[1] > 1 and A.stride(1) == 0)
    flags = 8 if ti else 0
    if lib.tlpc_lpc_summary_bytes(dt, B, T, N, flags, strategy, chunk_len) == 0:
        return 0
    fwd = lib.tlpc_lpc_forward_workspace(dt, B, T, N, flags, strategy, chunk_len)
    bwd = lib.tlpc_lpc_backward_workspace(dt, B, T, N, flags | 7, strategy, chunk_len)
    return max(fwd, bwd)


def recompute_supported(x: torch.Tensor, A: torch.Tensor, strategy: int = 0, chunk_len: int = 0) -> bool:
    """Whether the backward can rebuild y from the forward's chunk carries for these
    inputs (``tlpc_lpc_recompute_supported``, host only)."""
    import ctypes

    if x.dtype not in (torch.float32, torch.float64) or A.dim() != 3 or (A.shape[1] > 1 and A.stride(1) == 0):
        return False
    lib = _hiplib()
    i64, i32 = ctypes.c_int64, ctypes.c_int32
    fn = lib.tlpc_lpc_recompute_supported
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    B, T = x.shape
    return fn(_DTYPE_CODE[x.dtype], B, T, A.shape[-1], 7, strategy, chunk_len) == 1


def ragged_workspace_bytes(dtype: torch.dtype, G: int, N: int) -> int:
    """``tlpc_lpc_ragged_workspace`` for one order group of a ragged plan (host only)."""
    import ctypes

    fn = _hiplib().tlpc_lpc_ragged_workspace
    fn.restype = ctypes.c_size_t
    fn.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int32]
    return int(fn(_DTYPE_CODE[dtype], G, N))


def forward_workspace_bytes(x: torch.Tensor, A: torch.Tensor, strategy: int = 0, chunk_len: int = 0) -> int:
    """``tlpc_lpc_forward_workspace`` for these inputs (host only)."""
    import ctypes

    lib = _hiplib()
    i64, i32 = ctypes.c_int64, ctypes.c_int32
    fn = lib.tlpc_lpc_forward_workspace
    fn.restype = ctypes.c_size_t
    fn.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32]
    B, T = x.shape
    ti = A.dim() == 2 or (A.shape[1] > 1 and A.stride(1) == 0)
    return int(fn(_DTYPE_CODE[x.dtype], B, T, A.shape[-1], 8 if ti else 0, strategy, chunk_len))


def plan_of(x: torch.Tensor, A: torch.Tensor, strategy: int = 0, chunk_len: int = 0,
            backward: bool = False) -> tuple[int, int]:
    """(strategy, chunk length) that the forward (or backward) runs for these
    inputs: the C ABI's host-side plan query ``tlpc_lpc_plan``, which launches
    nothing. Strategy codes: 1 chunked, 2 per-signal, 3 serial."""
    import ctypes

    lib = _hiplib()
    i64, i32 = ctypes.c_int64, ctypes.c_int32
    fn = lib.tlpc_lpc_plan
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, i64, i64, i32, i32, i32, i32, i32,
                   ctypes.POINTER(i32), ctypes.POINTER(i32)]
    dt = _DTYPE_CODE[x.dtype]
    B, T = x.shape
    N = A.shape[-1]
    ti = A.dim() == 2 or (A.shape[1] > 1 and A.stride(1) == 0)
    flags = (8 if ti else 0) | (7 if backward else 0)
    s_out, l_out = i32(0), i32(0)
    rc = fn(dt, B, T, N, flags, strategy, chunk_len, int(backward), ctypes.byref(s_out), ctypes.byref(l_out))
    if rc != 0:
        raise RuntimeError(f"tlpc_lpc_plan failed ({rc})")
    return int(s_out.value), int(l_out.value)


def ops():
    """The ``torch.ops.torchlpc_amd`` namespace; raises if the library is missing."""
    if not EXTENSION_LOADED:
        raise RuntimeError(f"torchlpc_amd native extension is not loaded: {LOAD_ERROR}")
    return torch.ops.torchlpc_amd


def hip_runtimes_mapped() -> list[str]:
    """Paths of every libamdhip64 mapped into this process (expected: exactly one)."""
    try:
        maps = Path("/proc/self/maps").read_text()
    except OSError:  # pragma: no cover
        return []
    return sorted({ln.split()[-1] for ln in maps.splitlines() if "libamdhip64" in ln})


def _register_fake() -> None:
    """Shape functions (Meta / FakeTensor) so the ops trace under torch.compile."""
    lib = "torchlpc_amd::"

    @torch.library.register_fake(lib + "lpc")
    def _lpc_fake(x, A, zi, strategy=0, chunk_len=0):
        return torch.empty_like(x)

    @torch.library.register_fake(lib + "lpc_keep")
    def _lpc_keep_fake(x, A, zi, strategy=0, chunk_len=0):
        nbytes = keep_workspace_bytes(x, A, strategy, chunk_len)
        return torch.empty_like(x), x.new_empty((nbytes,), dtype=torch.uint8)

    @torch.library.register_fake(lib + "lpc_carries")
    def _lpc_carries_fake(x, A, zi, strategy=0, chunk_len=0):
        nbytes = forward_workspace_bytes(x, A, strategy, chunk_len)
        return torch.empty_like(x), x.new_empty((nbytes,), dtype=torch.uint8)

    @torch.library.register_fake(lib + "lpc_backward_recompute")
    def _lpc_bwd_rc_fake(g, A, zi, x, fws, need_x, need_a, need_zi, strategy=0, chunk_len=0):
        dx = torch.empty_like(g) if need_x else g.new_empty((0,))
        dA = g.new_empty(A.shape) if need_a else g.new_empty((0,))
        dzi = g.new_empty(zi.shape) if need_zi else g.new_empty((0,))
        return dx, dA, dzi

    @torch.library.register_fake(lib + "lpc_backward")
    def _lpc_bwd_fake(g, A, zi, y, need_x, need_a, need_zi, strategy=0, chunk_len=0, summary=None):
        B, T = g.shape
        N = A.shape[-1]
        dx = torch.empty_like(g) if need_x else g.new_empty((0,))
        if need_a:
            dA = g.new_empty((B, T, N)) if A.dim() == 3 else g.new_empty((B, N))
        else:
            dA = g.new_empty((0,))
        dzi = g.new_empty((B, N)) if need_zi else g.new_empty((0,))
        return dx, dA, dzi

    @torch.library.register_fake(lib + "lpc_ragged_")
    def _lpc_ragged_fake(x, A, zi, table, y, K, G, max_chunks, N, chunk_len, extents):
        return None

    @torch.library.register_fake(lib + "lpc_ragged_keep_")
    def _lpc_ragged_keep_fake(x, A, zi, table, y, K, G, max_chunks, N, chunk_len, extents):
        return x.new_empty((ragged_workspace_bytes(x.dtype, G, N),), dtype=torch.uint8)

    @torch.library.register_fake(lib + "lpc_ragged_backward_")
    def _lpc_ragged_bwd_fake(g, A, zi, y, table, dx, dA, dzi, K, G, max_chunks, N, chunk_len, nx, na, nz, extents,
                             kept=None):
        return None

    @torch.library.register_fake(lib + "scan")
    def _scan_fake(x, w, init, chunk_len=0):
        return torch.empty_like(x)

    @torch.library.register_fake(lib + "scan_backward")
    def _scan_bwd_fake(g, w, init, h, need_x, need_w, need_init, chunk_len=0):
        dx = torch.empty_like(g) if need_x else g.new_empty((0,))
        dw = torch.empty_like(g) if need_w else g.new_empty((0,))
        dinit = g.new_empty((g.shape[0],)) if need_init else g.new_empty((0,))
        return dx, dw, dinit


_load()
