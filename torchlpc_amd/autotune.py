"""Project-written autotuner for the filter launches.

The kernels take two launch parameters:

* the strategy: the per-signal tiled kernels (one workgroup, 8 waves, per
  signal; float32 order 16), the chunked multi-kernel scan, or serial;
* the chunk length of the chunked scan (samples per independently scanned
  chunk). It sets the parallelism inside a signal: the order-16 reruns run one
  4-lane quad per chunk, i.e. one wave per 16 chunks, so a signal of T samples
  gets ceil(T / chunk / 16) rerun waves, and the carry chain one link per chunk.

``tune()`` times candidate (strategy, chunk length) pairs on the caller's real
tensors with HIP events, as the whole step a training call runs (forward, then
the fused backward with the forward's plan and kept workspace), and stores the
fastest per key (op, dtype, device, B, T, N, time-invariant); ``lookup()``
returns the stored choice or ``(0, 0)`` (the library heuristic). The table is
kept in memory and persisted as JSON (``TORCHLPC_AMD_TUNE_CACHE``, default
``~/.cache/torchlpc_amd/tune.json``). Lookups are read-mostly under a lock and
never synchronise the device, so they are safe inside hipGraph capture;
``tune()`` synchronises and must run outside capture.

The reference has no tuner: its chunking constants are hard-coded
(``torchlpc/csrc/cuda/linear_recurrence.cu:238-239``, ``torchlpc/recurrence.py:49,64``).
"""
from __future__ import annotations

import json
import os
import threading
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Tuple

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
    dev = torch.cuda.get_device_name(x.device) if x.is_cuda else "cpu"
    return f"{op}|{x.dtype}|{dev}|B{B}|T{T}|N{N}|ti{int(ti)}"


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


def default_candidates(B: int, T: int, N: int, dtype: torch.dtype, ti: bool = False) -> List[Tuple[int, int]]:
    """(strategy, chunk_len) pairs worth timing for a batch shape.

    The heuristic (0, 0) always; the per-signal kernels for float32 order 16
    (time-varying) and orders 16..32 (time-invariant a[B, N]); the chunked scan
    at chunk lengths that give between 1/4 and 4 times as many rerun lanes as
    compute units can hold (chunk lengths are multiples of 16, 32 from order 17
    up)."""
    blk = 32 if N > 16 else 16
    out: List[Tuple[int, int]] = [(STRATEGY_AUTO, 0)]
    if dtype == torch.float32 and (N == 16 or (ti and 16 <= N <= 32)):
        out.append((STRATEGY_SIGNAL, 0))
    if T <= 0:
        return out
    for chunks_per_signal in (16, 32, 64, 128, 256):
        L = -(-T // chunks_per_signal)
        L = max(blk, -(-L // blk) * blk)
        if L <= T and (STRATEGY_CHUNKED, L) not in out:
            out.append((STRATEGY_CHUNKED, L))
    return out


def _time(fn, reps: int) -> float:
    """Median per-call milliseconds of `fn` over `reps` event-timed calls."""
    fn()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        fn()
        end.record()
        torch.cuda.synchronize()
        times.append(start.elapsed_time(end))
    times.sort()
    return times[len(times) // 2]


def step_fn(x: torch.Tensor, A: torch.Tensor, zi: torch.Tensor, grad: Optional[torch.Tensor],
            choice: Tuple[int, int]):
    """The call sequence one training step runs for a (strategy, chunk) choice:
    the forward keeping its workspace, then (with `grad`) the fused backward with
    the same plan, reusing the kept summaries when there are any."""
    from ._ext import ops

    s, L = choice

    def run():
        if grad is None:
            return ops().lpc(x, A, zi, s, L)
        y, ws = ops().lpc_keep(x, A, zi, s, L)
        return y, ops().lpc_backward(grad, A, zi, y, True, True, True, s, L, ws if ws.numel() else None)

    return run


def tune(
    x: torch.Tensor,
    A: torch.Tensor,
    zi: torch.Tensor,
    grad: Optional[torch.Tensor] = None,
    candidates: Optional[Iterable[Tuple[int, int]]] = None,
    reps: int = 5,
    persist: bool = True,
) -> Dict[str, object]:
    """Time the step (forward, and with ``grad`` the fused backward) for every
    candidate and store the fastest for ``lpc_forward`` and ``lpc_backward``.

    Returns ``{"choice": (s, L), "ms": {"s,L": median ms, ...},
    "heuristic_ms": ms of (0, 0)}``. Must not be called during hipGraph capture
    (it synchronises).
    """
    assert x.is_cuda, "tuning needs GPU tensors"
    B, T = x.shape
    N = A.shape[-1]
    ti = A.dim() == 2 or (A.dim() == 3 and A.shape[1] > 1 and A.stride(1) == 0)
    cands = list(candidates) if candidates is not None else default_candidates(B, T, N, x.dtype, ti)
    if (STRATEGY_AUTO, 0) not in cands:
        cands.insert(0, (STRATEGY_AUTO, 0))
    ms = {c: _time(step_fn(x, A, zi, grad, c), reps) for c in cands}
    best = min(cands, key=lambda c: ms[c])
    store("lpc_forward", x, A, best, persist)
    if grad is not None:
        store("lpc_backward", grad, A, best, persist)
    return {
        "choice": best,
        "ms": {f"{c[0]},{c[1]}": v for c, v in ms.items()},
        "heuristic_ms": ms[(STRATEGY_AUTO, 0)],
    }
