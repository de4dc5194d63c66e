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
    EXTENSION_LOADED = True
    _register_fake()


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

    @torch.library.register_fake(lib + "lpc_backward")
    def _lpc_bwd_fake(g, A, zi, y, need_x, need_a, need_zi, strategy=0, chunk_len=0):
        B, T = g.shape
        N = A.shape[-1]
        dx = torch.empty_like(g) if need_x else g.new_empty((0,))
        if need_a:
            dA = g.new_empty((B, T, N)) if A.dim() == 3 else g.new_empty((B, N))
        else:
            dA = g.new_empty((0,))
        dzi = g.new_empty((B, N)) if need_zi else g.new_empty((0,))
        return dx, dA, dzi

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
