"""Project-written autotuner for the filter launches.

The kernels take two launch parameters: the strategy (chunked multi-kernel scan,
one workgroup per signal, or serial) and the chunk length (samples per
independently scanned chunk). ``tune()`` times the candidates on the caller's
real tensors with HIP events and stores the winner per key
(op, dtype, device, B, log2-bucket of T, N, time-invariant); ``lookup()`` returns
the stored choice or ``(0, 0)`` (library heuristic). The table is kept in memory
and persisted as JSON (``TORCHLPC_AMD_TUNE_CACHE``, default
``~/.cache/torchlpc_amd/tune.json``). Lookups are read-mostly under a lock and
never synchronise the device, so they are safe inside hipGraph capture.

The reference has no tuner: its chunking constants are hard-coded
(``torchlpc/csrc/cuda/linear_recurrence.cu:238-239``, ``torchlpc/recurrence.py:49,64``).
"""
from __future__ import annotations

import json
import math
import os
import threading
from pathlib import Path
from typing import Dict, Iterable, Optional, Tuple

import torch

STRATEGY_AUTO = 0
STRATEGY_CHUNKED = 1
STRATEGY_SIGNAL = 2
STRATEGY_SERIAL = 3

_lock = threading.Lock()
_table: Dict[str, Tuple[int, int]] = {}
_loaded = False


def cache_path() -> Path:
    p = os.environ.get("TORCHLPC_AMD_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "torchlpc_amd" / "tune.json"


def _key(op: str, x: torch.Tensor, A: torch.Tensor) -> str:
    B, T = x.shape[0], x.shape[1]
    N = A.shape[-1]
    ti = A.dim() == 2 or (A.dim() == 3 and A.shape[1] > 1 and A.stride(1) == 0)
    tb = int(math.log2(T)) if T > 0 else -1
    dev = torch.cuda.get_device_name(x.device) if x.is_cuda else "cpu"
    return f"{op}|{x.dtype}|{dev}|B{B}|T{tb}|N{N}|ti{int(ti)}"


def _load() -> None:
    global _loaded
    if _loaded:
        return
    with _lock:
        if _loaded:
            return
        try:
            data = json.loads(cache_path().read_text())
            for k, v in data.items():
                _table[k] = (int(v[0]), int(v[1]))
        except (OSError, ValueError):
            pass
        _loaded = True


def lookup(op: str, x: torch.Tensor, A: torch.Tensor) -> Tuple[int, int]:
    """Tuned (strategy, chunk_len) for this launch, or (0, 0)."""
    if not _table and _loaded:
        return (0, 0)
    _load()
    if not _table:
        return (0, 0)
    with _lock:
        return _table.get(_key(op, x, A), (0, 0))


def store(op: str, x: torch.Tensor, A: torch.Tensor, choice: Tuple[int, int], persist: bool = True) -> None:
    _load()
    with _lock:
        _table[_key(op, x, A)] = (int(choice[0]), int(choice[1]))
        if persist:
            path = cache_path()
            try:
                path.parent.mkdir(parents=True, exist_ok=True)
                tmp = path.with_suffix(".tmp")
                tmp.write_text(json.dumps({k: list(v) for k, v in _table.items()}, indent=1))
                os.replace(tmp, path)
            except OSError:
                pass


def clear() -> None:
    global _loaded
    with _lock:
        _table.clear()
        _loaded = True


def default_candidates(T: int, N: int) -> Iterable[Tuple[int, int]]:
    """(strategy, chunk_len) pairs worth timing for a signal length and order."""
    blk = 32 if N > 16 else 16
    out = [(STRATEGY_AUTO, 0)]
    if N == 16:
        out.append((STRATEGY_SIGNAL, 0))
    L = max(blk, 64)
    while L <= min(8192, max(T, blk)):
        out.append((STRATEGY_CHUNKED, L))
        L *= 2
    return out


def _time(fn, reps: int) -> float:
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps


def tune(
    x: torch.Tensor,
    A: torch.Tensor,
    zi: torch.Tensor,
    grad: Optional[torch.Tensor] = None,
    candidates: Optional[Iterable[Tuple[int, int]]] = None,
    reps: int = 5,
    persist: bool = True,
) -> Dict[str, Tuple[int, int]]:
    """Time forward (and, with ``grad``, the fused backward) over the candidates.

    Must not be called during hipGraph capture (it synchronises).
    """
    from ._ext import ops

    assert x.is_cuda, "tuning needs GPU tensors"
    N = A.shape[-1]
    cands = list(candidates) if candidates is not None else list(default_candidates(x.shape[1], N))
    best_f = min(cands, key=lambda c: _time(lambda: ops().lpc(x, A, zi, c[0], c[1]), reps))
    store("lpc_forward", x, A, best_f, persist)
    result = {"lpc_forward": best_f}
    if grad is not None:
        y = ops().lpc(x, A, zi, best_f[0], best_f[1])
        best_b = min(
            cands,
            key=lambda c: _time(
                lambda: ops().lpc_backward(grad, A, zi, y, True, True, True, c[0], c[1]), reps
            ),
        )
        store("lpc_backward", grad, A, best_b, persist)
        result["lpc_backward"] = best_b
    return result
