#!/usr/bin/env python3
"""Benchmark of the all-pole (LPC) filter hot path on one MI355X.

Workload (one step): forward filter + fused adjoint (dx, dA, dzi) at B=256
signals, T=48000 samples, order N=16, float32, time-varying coefficients from
interpolated reflection coefficients (anchors every 240 samples, step-up
recursion), x ~ N(0,1), zi ~ N(0, 0.1^2), upstream gradient w ~ N(0,1)
(loss = sum(y * w)). Inputs are generated on the GPU from a fixed seed and are
resident in HBM before timing starts.

Prints one JSON line: Msamples/s over the whole step, the HBM roofline of the
dominant op measured with HIP events on its stream, the plain GPU form
(forward kernel + unfused, reference-style backward) and the reference CPU path
timed on this host's cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-plain]
  python bench.py --trace-summary <rocprofv3 results.db>
  python bench.py --pmc-summary <FETCH_SIZE run dir> <WRITE_SIZE run dir>

"Dominant kernel" in the roofline object is the dominant op of the step (the
backward at this shape): its algorithmic bytes per call over its event-timed
duration. The op is one kernel (per-signal strategy) or a short fixed sequence
of kernels (chunked strategy; OP_KERNELS_BY_STRATEGY); the rocprof trace summary
adds their average durations for comparison, and `traffic` is their
counter-measured HBM bytes per call from the committed summary. Before timing,
the project tuner picks the (strategy, chunk length) plan on the same inputs
(--no-tune: the library heuristic).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published peak
COPY_BYTES = 1 << 30  # the --calibrate-copy kernel reads and writes this many bytes
PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")

# kernels of one forward and one backward call at the order-16 float32 shape, per
# strategy (name prefixes: template arguments after these select the load variant)
OP_KERNELS_BY_STRATEGY = {
    2: {  # per-signal tiled kernels
        "lpc (forward)": ("tile_fwd",),
        "lpc_backward (fused adjoint)": ("tile_bwd",),
    },
    1: {  # chunked scan
        "lpc (forward)": ("sum_fwd<true", "carry16<true>", "rerun_fwd"),
        "lpc_backward (fused adjoint)": ("sum_adj", "carry16<false>", "rerun_adj<true"),
    },
}


def _op_of(avg, prefix):
    """Total of the kernels in `avg` whose name starts with `prefix` (None if none)."""
    hits = [v for k, v in avg.items() if k.startswith(prefix)]
    return sum(hits) if hits else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(B, T, N, device, seed=1002, hop=240, kmax=0.9, zi_std=0.1):
    """Seeded synthetic inputs for the B=256, T=48000, N=16 workload (FP64, cast to FP32)."""
    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    f64 = dict(device=device, dtype=torch.float64)
    x = torch.randn(B, T, generator=g, **f64)
    n_anchor = (T - 1) // hop + 2
    k_anchor = (torch.rand(B, n_anchor, N, generator=g, **f64) * 2 - 1) * kmax
    pos = torch.arange(T, **f64) / hop
    lo = pos.floor().long().clamp(max=n_anchor - 1)
    hi = (lo + 1).clamp(max=n_anchor - 1)
    frac = (pos - lo)[None, :, None]
    k = k_anchor[:, lo] * (1 - frac) + k_anchor[:, hi] * frac
    del k_anchor
    a = torch.zeros_like(k)
    for m in range(1, N + 1):  # step-up recursion: a stable all-pole filter per sample
        km = k[..., m - 1 : m]
        prev = a[..., : m - 1].clone()
        a[..., : m - 1] = prev + km * prev.flip(-1)
        a[..., m - 1] = k[..., m - 1]
    del k
    zi = torch.randn(B, N, generator=g, **f64) * zi_std
    w = torch.randn(B, T, generator=g, **f64)
    f32 = torch.float32
    return x.to(f32), a.to(f32).contiguous(), zi.to(f32), w.to(f32)


def cpu_baseline(T, N, sample_signals=64, min_seconds=10.0):
    """Reference CPU path (torchlpc/csrc/scan_cpu.cpp op + the core.py:186-229 backward
    composition), or the C restatement when the reference op is not built. The
    sample is `sample_signals` signals of the workload, run repeatedly for about
    `min_seconds` of CPU time (median run)."""
    import numpy as np
    import torch

    from oracle import oracle as O

    cores = os.cpu_count() or 1
    Bs = sample_signals
    rng = np.random.default_rng(7)
    x, A, zi, w = O.make_reflection_inputs(Bs, T, N, 240, 7)
    x, A, zi, w = (v.astype(np.float32) for v in (x, A, zi, w))
    if O.REF_LIB_PATH.exists():
        kind = "reference"
        torch.ops.load_library(str(O.REF_LIB_PATH))
        torch.set_num_threads(cores)
        tx, tA, tzi, tw = (torch.from_numpy(np.ascontiguousarray(v)) for v in (x, A, zi, w))

        def run():
            y = torch.ops.torchlpc.lpc(tx, tA, tzi)
            O.reference_backward_torch(tw, tA, tzi, y)

        threads = torch.get_num_threads()
    else:
        kind = "port"
        O.set_threads(cores)

        def run():
            y = O.lpc(x, A, zi)
            O.lpc_backward(w, A, zi, y)

        threads = O.max_threads()
    del rng
    run()
    times = []
    t_all = time.perf_counter()
    while len(times) < 3 or time.perf_counter() - t_all < min_seconds:
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    reps = len(times)
    return {
        "value": Bs * T / t / 1e6,
        "unit": "Msamples/s",
        "cores": int(min(threads, Bs)),
        "kind": kind,
        "sample": f"{Bs} signals x T={T}, N={N}, fp32, fwd+bwd, median of {reps} after 1 warm-up "
        f"(parallel over signals only, so effective parallelism = min(signals, threads) = "
        f"{min(threads, Bs)}; host os.cpu_count()={cores})",
        "seconds_per_run": t,
        "cpu_model": _cpu_model(),
    }


def accuracy_sample(x, A, zi, w, outs, signals):
    """Relative L2 errors of one timed step's outputs against the FP64 oracle, on a
    subset of the batch's signals (signals are independent, so the rows of the
    full-batch outputs are checked as they are)."""
    import numpy as np

    from oracle import oracle as O

    idx = list(signals)
    h = lambda t: t[idx].detach().cpu().numpy().astype(np.float64)  # noqa: E731
    y, (dx, dA, dzi) = outs
    x64, A64, zi64, w64 = h(x), h(A), h(zi), h(w)
    y64 = O.lpc(x64, A64, zi64)
    rx, rA, rzi = O.lpc_backward(w64, A64, zi64, y64)
    N = A64.shape[-1]
    yo = y[idx].detach().cpu().numpy()
    return {
        "y": O.rel_err(yo, y64),
        "zf": O.rel_err(yo[:, ::-1][:, :N], y64[:, ::-1][:, :N]),
        "dx": O.rel_err(h(dx), rx),
        "dA": O.rel_err(h(dA), rA),
        "dzi": O.rel_err(h(dzi), rzi),
        "signals": idx,
        "bar": 1e-5,
        "reference": "FP64 oracle (C restatement of the reference loop) on the float32 inputs",
    }


def _cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _step_up(k):
    """Direct-form coefficients a[..., N] from reflection coefficients k[..., N]."""
    import torch

    a = torch.zeros_like(k)
    N = k.shape[-1]
    for m in range(1, N + 1):
        km = k[..., m - 1 : m]
        prev = a[..., : m - 1].clone()
        a[..., : m - 1] = prev + km * prev.flip(-1)
        a[..., m - 1] = k[..., m - 1]
    return a


def _interp(anchors, T, hop):
    """Linear interpolation of anchor values [..., n_anchor, N] every `hop` samples."""
    import torch

    n_anchor = anchors.shape[-2]
    pos = torch.arange(T, device=anchors.device, dtype=anchors.dtype) / hop
    lo = pos.floor().long().clamp(max=n_anchor - 1)
    hi = (lo + 1).clamp(max=n_anchor - 1)
    frac = (pos - lo)[:, None]
    return anchors[..., lo, :] * (1 - frac) + anchors[..., hi, :] * frac


def ragged512_inputs(device, K=512, seed=1004, hop=256):
    """K ragged signals: lengths ~ U{1024..65536}, orders uniform over {2,4,8,16,32},
    packed in order of their filter order (all order-2 signals first, and so on),
    so that each order's samples, coefficient rows and states are one contiguous
    slice. Coefficients come from reflection coefficients k ~ U(-0.9, 0.9) at
    anchors every `hop` samples (U(-0.5, 0.5) at order 32, where 0.9 makes the
    time-varying filter grow past the float32 range), interpolated per sample and
    converted by the step-up recursion. Each order is generated in a few batched
    device ops (about 20 plus 5 per order of the step-up), not per signal.
    Returns (lengths, orders, x, A, zi) with float64 device tensors."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(seed)
    lengths = torch.randint(1024, 65537, (K,), generator=g)
    orders = torch.tensor((2, 4, 8, 16, 32))[torch.randint(0, 5, (K,), generator=g)]
    perm = torch.sort(orders, stable=True).indices
    lengths, orders = lengths[perm], orders[perm]
    gd = torch.Generator(device=device).manual_seed(seed)
    f64 = dict(device=device, dtype=torch.float64)
    x = torch.randn(int(lengths.sum()), generator=gd, **f64)
    zi = torch.randn(int(orders.sum()), generator=gd, **f64) * 0.1
    parts = []
    for n in (2, 4, 8, 16, 32):
        sel = (orders == n).nonzero().flatten()
        if sel.numel() == 0:
            continue
        lens = lengths[sel].to(device)
        nanc = (lens - 1) // hop + 2
        kmax = 0.9 if n <= 16 else 0.5
        anchors = (torch.rand(int(nanc.sum()), n, generator=gd, **f64) * 2 - 1) * kmax
        zero = torch.zeros(1, dtype=torch.int64, device=device)
        abase = torch.cat([zero, nanc.cumsum(0)[:-1]])
        start = torch.cat([zero, lens.cumsum(0)[:-1]])
        sig = torch.repeat_interleave(torch.arange(sel.numel(), device=device), lens)
        t = torch.arange(sig.numel(), device=device) - start[sig]
        pos = t.to(torch.float64) / hop
        last = nanc[sig] - 1
        lo = torch.minimum(pos.floor().long(), last)
        hi = torch.minimum(lo + 1, last)
        frac = (pos - lo)[:, None]
        k = anchors[abase[sig] + lo] * (1 - frac) + anchors[abase[sig] + hi] * frac
        del anchors, sig, t, pos, lo, hi, frac
        parts.append(_step_up(k).reshape(-1))
        del k
    return lengths.tolist(), orders.tolist(), x, torch.cat(parts), zi


def workload_inputs(name, device, seed):
    """Synthetic inputs of the secondary workloads (float64, generated on the GPU)."""
    import math

    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    f64 = dict(device=device, dtype=torch.float64)
    if name == "b16-t8192-n4":  # B=16, T=8192, N=4: two resonators per signal, 32 anchor frames
        B, T = 16, 8192
        r = torch.rand(B, 32, 2, generator=g, **f64) * 0.45 + 0.5
        th = (torch.rand(B, 32, 2, generator=g, **f64) * 0.9 + 0.05) * math.pi
        rt, tt = _interp(r, T, (T - 1) / 31), _interp(th, T, (T - 1) / 31)
        c1, c2 = -2 * rt * torch.cos(tt), rt * rt  # each section 1 + c1 z^-1 + c2 z^-2
        A = torch.stack(
            [
                c1[..., 0] + c1[..., 1],
                c2[..., 0] + c2[..., 1] + c1[..., 0] * c1[..., 1],
                c1[..., 0] * c2[..., 1] + c2[..., 0] * c1[..., 1],
                c2[..., 0] * c2[..., 1],
            ],
            -1,
        )
        return torch.randn(B, T, generator=g, **f64), A, torch.zeros(B, 4, **f64), None
    if name == "b4-t1048576-n2":  # B=4, T=2^20, N=2: slowly drifting resonators near the unit circle
        B, T = 4, 1 << 20
        r = torch.rand(B, 1, generator=g, **f64) * 0.099 + 0.9
        thc = (torch.rand(B, 1, generator=g, **f64) * 0.6 + 0.2) * math.pi
        phi = torch.rand(B, 1, generator=g, **f64) * 2 * math.pi
        t = torch.arange(T, **f64)[None]
        th = thc + 0.1 * math.pi * torch.sin(2 * math.pi * t / 65536 + phi)
        A = torch.stack([-2 * r * torch.cos(th), (r * r).expand(B, T)], -1)
        return (torch.randn(B, T, generator=g, **f64), A, torch.randn(B, 2, generator=g, **f64),
                torch.randn(B, T, generator=g, **f64))
    if name in ("ti-b1024-t16384-n32", "ti-b4-t1048576-n32"):  # time-invariant a[B, N], N=32
        B, T, N = (1024, 16384, 32) if name == "ti-b1024-t16384-n32" else (4, 1 << 20, 32)
        a = _step_up((torch.rand(B, N, generator=g, **f64) * 2 - 1) * 0.8)
        return (torch.randn(B, T, generator=g, **f64), a, torch.randn(B, N, generator=g, **f64) * 0.1,
                torch.randn(B, T, generator=g, **f64))
    raise ValueError(name)


def bench_workload(name, steps, warmup, ragged_order=0, tune=True):
    """fwd(+bwd) of a secondary workload; returns the JSON record. For the ragged
    workload, ragged_order > 0 runs only that order's group of signals (on the
    caller's stream), for a kernel trace of one group."""
    import torch

    import torchlpc_amd

    dev = torch.device("cuda:0")
    ops = torch.ops.torchlpc_amd
    stream = torch.cuda.current_stream()
    f32 = torch.float32
    if name == "ragged512":  # 512 ragged signals, lengths ~ U{1024..65536}, orders in {2,4,8,16,32}
        K = 512
        lengths, orders, x, A, zi = ragged512_inputs(dev, K)
        x, A, zi = x.to(f32), A.to(f32), zi.to(f32)
        gd = torch.Generator(device=dev).manual_seed(1014)
        w = torch.randn(x.numel(), generator=gd, device=dev, dtype=f32)
        plan = torchlpc_amd.RaggedPlan(lengths, orders, dev, f32)
        heuristic_chunks = {grp[0]: grp[5] for grp in plan.groups}
        # untimed: the project tuner picks each order group's chunk length on these inputs
        tuned = plan.tune(x, A, zi, w) if tune else None
        y = torch.empty_like(x)
        dx, dA, dzi = torch.empty_like(x), torch.empty_like(A), torch.empty_like(zi)

        from torchlpc_amd.ragged import keeps_summaries

        kept = {}

        # as the autograd Function runs a training step: the float32 order-16
        # and order-32 groups keep their forward's summaries for the backward
        def fwd(gi, grp):
            n, table, Kg, G, maxp, L, ext = grp
            if keeps_summaries(grp, f32):
                kept[gi] = ops.lpc_ragged_keep_(x, A, zi, table, y, Kg, G, maxp, n, L, ext)
            else:
                ops.lpc_ragged_(x, A, zi, table, y, Kg, G, maxp, n, L, ext)

        def bwd(gi, grp):
            n, table, Kg, G, maxp, L, ext = grp
            ops.lpc_ragged_backward_(w, A, zi, y, table, dx, dA, dzi, Kg, G, maxp, n, L, True, True, True, ext,
                                     kept.pop(gi, None))

        if ragged_order and ragged_order not in orders:
            raise SystemExit(f"--ragged-order {ragged_order}: no signal of that order (orders are 2, 4, 8, 16, 32)")

        def step():
            if ragged_order:
                for gi, grp in enumerate(plan.groups):
                    if grp[0] == ragged_order:
                        fwd(gi, grp)
                        bwd(gi, grp)
                return
            # the order groups, one after another on the current stream
            plan.run_indexed(fwd)
            plan.run_indexed(bwd)

        sel = [(T, N) for T, N in zip(lengths, orders) if not ragged_order or N == ragged_order]
        S = sum(T for T, _ in sel)
        alg = sum((3 * N + 5) * 4 * T for T, N in sel)
        workload = (f"ragged all-pole filter fwd + fused adjoint bwd, {K} signals, T_k~U{{1024..65536}}, "
                    "N_k in {2,4,8,16,32}, fp32, one HIP graph; reflection coefficients k~U(-0.9,0.9) "
                    "at orders up to 16 and k~U(-0.5,0.5) at order 32 (0.9 overflows float32 there)")
        if ragged_order:
            workload += f", only the {len(sel)} signals of order {ragged_order}"
        cfg = {"workload": workload, "signals": len(sel), "samples": S,
               "chunk_len_by_order": {grp[0]: grp[5] for grp in plan.groups},
               "heuristic_chunk_len_by_order": heuristic_chunks,
               "tuned": {n: {"choice": v["choice"], "ms": v["ms"]} for n, v in tuned.items()} if tuned else None}
        s = torch.cuda.Stream()
        s.wait_stream(stream)
        with torch.cuda.stream(s):
            step()
        stream.wait_stream(s)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            step()
        run = graph.replay
    else:
        x64, A64, zi64, w64 = workload_inputs(name, dev, {"b16-t8192-n4": 1001, "b4-t1048576-n2": 1003, "ti-b1024-t16384-n32": 1005, "ti-b4-t1048576-n32": 1006}[name])
        x, A, zi = x64.to(f32), A64.to(f32).contiguous(), zi64.to(f32)
        w = w64.to(f32) if w64 is not None else None
        B, T = x.shape
        N = A.shape[-1]
        S = B * T
        if name == "b16-t8192-n4":
            def run():
                ops.lpc(x, A, zi)
            alg = (N + 2) * 4 * S
            cfg = {"workload": f"time-varying all-pole filter forward, B={B}, T={T}, N={N}, fp32", "B": B, "T": T, "N": N}
        else:
            def run():
                y = ops.lpc(x, A, zi)
                ops.lpc_backward(w, A, zi, y, True, True, True)
            ti = A.dim() == 2
            alg = (5 if ti else 3 * N + 5) * 4 * S
            kind = "time-invariant" if ti else "time-varying"
            cfg = {"workload": f"{kind} all-pole filter fwd + fused adjoint bwd, B={B}, T={T}, N={N}, fp32", "B": B, "T": T, "N": N}
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    gbs = alg / (ms * 1e-3) / 1e9
    extra = {}
    if name == "b4-t1048576-n2":
        # the same step with the adjoint that rebuilds y from the forward's chunk
        # carries (what the autograd Function runs at this length): time, and the
        # bytes kept between the passes instead of y
        def run_rc():
            y, ws = ops.lpc_carries(x, A, zi, 0, 0)
            ops.lpc_backward_recompute(w, A, zi, x, ws, True, True, True, 0, 0)
            return ws

        for _ in range(warmup):
            ws = run_rc()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(steps):
            ws = run_rc()
        e1.record(stream)
        torch.cuda.synchronize()
        ms_rc = e0.elapsed_time(e1) / steps
        extra["rebuild_y"] = {
            "ms_per_step": ms_rc,
            "Msamples_per_s": S / (ms_rc * 1e-3) / 1e6,
            "kept_bytes": int(ws.numel()),
            "y_bytes": int(x.numel() * x.element_size()),
            "memory_saved_bytes": int(x.numel() * x.element_size() - ws.numel()),
        }
    return {
        "metric": "Msamples/s (secondary workload)",
        "value": S / (ms * 1e-3) / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": alg},
        **extra,
    }


def _short(name):
    """Kernel name without namespace, return type and argument list."""
    name = name.replace("(anonymous namespace)::", "")
    name = name.split("(")[0].replace("void ", "").strip()
    return name.split("::")[-1]


def trace_summary(db_path):
    """Per-kernel average durations of a rocprofv3 --kernel-trace database, and the
    per-call total of the kernels of each op."""
    import sqlite3

    con = sqlite3.connect(db_path)
    rows = con.execute(
        "select name, count(*), avg(duration) from kernels group by name order by sum(duration) desc"
    ).fetchall()
    lines = [f"{'kernel':40s} {'calls':>5s} {'avg_us':>10s}"]
    avg = {}
    for name, n, d in rows:
        k = _short(name)
        avg[k] = d / 1e3
        if "tlpc" in name or k.startswith("neg"):
            lines.append(f"{k:40s} {n:5d} {d / 1e3:10.1f}")
    for kernels in OP_KERNELS_BY_STRATEGY.values():
        for op, ks in kernels.items():
            parts = [_op_of(avg, k) for k in ks]
            if all(p is not None for p in parts):
                lines.append(f"op {op}: sum of kernel averages = {sum(parts):.1f} us ({' + '.join(ks)})")
    return "\n".join(lines)


def pmc_summary(fetch_dir, write_dir):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) per kernel, scaled by what the same
    counters read for the --calibrate-copy kernel (COPY_BYTES read and written)."""
    import csv
    import glob

    ours = set()  # this library's kernels (namespace tlpc)

    def load(d, counter):
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == counter:
                    k = _short(r["Kernel_Name"])
                    if "tlpc::" in r["Kernel_Name"]:
                        ours.add(k)
                    vals.setdefault(k, []).append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in vals.items()}

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    copy = next((k for k in fetch if "neg" in k), None)
    if copy is None:
        raise SystemExit("no --calibrate-copy kernel in the counter runs")
    ff = COPY_BYTES / (fetch[copy] * 1024)
    fw = COPY_BYTES / (write[copy] * 1024)
    lines = [
        f"calibration kernel {copy}: FETCH_SIZE {fetch[copy]:.0f} KiB, WRITE_SIZE {write[copy]:.0f} KiB "
        f"for {COPY_BYTES} bytes each -> scale fetch x{ff:.3f}, write x{fw:.3f}",
        f"{'kernel':40s} {'FETCH_KiB':>12s} {'WRITE_KiB':>12s} {'traffic_MB':>11s}",
    ]
    traffic = {}
    for k in sorted(fetch):
        if k not in write:
            continue
        t = fetch[k] * 1024 * ff + write[k] * 1024 * fw
        traffic[k] = t
        if k in ours:
            lines.append(f"{k:40s} {fetch[k]:12.0f} {write[k]:12.0f} {t / 1e6:11.1f}")
    for strategy, kernels in OP_KERNELS_BY_STRATEGY.items():
        for op, ks in kernels.items():
            parts = [_op_of(traffic, k) for k in ks]
            if all(p is not None for p in parts):
                lines.append(f"traffic_bytes_per_call strategy{strategy} {op}: {sum(parts):.0f}")
    return "\n".join(lines)


def committed_traffic(op, B, T, N, strategy):
    """HBM bytes per call of `op` under `strategy` from the committed counter
    summary, if there is one for this shape."""
    path = os.path.join(PROFILES, f"lpc_traffic_B{B}_T{T}_N{N}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_call strategy{strategy} {op}:"):
            return float(line.rsplit(":", 1)[1])
    return None


def calibrate_copy():
    """One elementwise kernel that reads and writes COPY_BYTES, whose FETCH_SIZE /
    WRITE_SIZE readings scale the counters of the other kernels (--pmc-summary)."""
    import torch

    src = torch.ones(COPY_BYTES // 4, device="cuda:0")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    del src, dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--B", type=int, default=256)
    ap.add_argument("--T", type=int, default=48000)
    ap.add_argument("--N", type=int, default=16)
    ap.add_argument("--strategy", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="time the library heuristic instead of the tuned (strategy, chunk) choice")
    ap.add_argument("--ragged-order", type=int, default=0,
                    help="with --workload ragged512: run only the signals of this order (one group)")
    ap.add_argument("--workload", default="b256-t48000-n16", choices=["b256-t48000-n16", "b16-t8192-n4", "b4-t1048576-n2", "ragged512", "ti-b1024-t16384-n32", "ti-b4-t1048576-n32"],
                    help="the other workloads print their own line (the default is B=256, T=48000, N=16)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip timing the secondary workloads after the headline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after timing, run one copy-like kernel of known size (for counter scaling)")
    ap.add_argument("--trace-summary", metavar="DB", help="summarise a rocprofv3 kernel-trace database and exit")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise rocprofv3 FETCH_SIZE / WRITE_SIZE runs and exit")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        return
    if args.pmc_summary:
        print(pmc_summary(*args.pmc_summary))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(
            "bench.py measures one GPU: this build runs the filter on a single MI355X "
            "(signals shard across GPUs with no communication). "
            f"--gpus {args.gpus} is not supported.",
            file=sys.stderr,
        )
        sys.exit(2)

    if args.workload != "b256-t48000-n16":
        log(f"[bench] workload {args.workload}")
        print(json.dumps(bench_workload(args.workload, args.steps, args.warmup, args.ragged_order,
                                        tune=not args.no_tune)), flush=True)
        if args.calibrate_copy:
            calibrate_copy()
        return

    import torch

    import torchlpc_amd
    from torchlpc_amd import autotune
    from torchlpc_amd.core import composite_backward

    if not torchlpc_amd.EXTENSION_LOADED:
        raise RuntimeError(torchlpc_amd.LOAD_ERROR)
    dev = torch.device("cuda:0")
    B, T, N = args.B, args.T, args.N
    log(f"[bench] generating inputs B={B} T={T} N={N}")
    x, A, zi, w = make_inputs(B, T, N, dev)
    torch.cuda.synchronize()
    ops = torch.ops.torchlpc_amd
    tuned = None
    if args.strategy or args.chunk:
        s_f, L_f = args.strategy, args.chunk
    elif args.no_tune:
        s_f, L_f = 0, 0
    else:
        # untimed: the project tuner times the candidate plans on these inputs
        # (the whole step each) and the timed steps below run its choice
        log("[bench] tuning (strategy, chunk length)")
        tuned = autotune.tune(x, A, zi, grad=w, reps=3, persist=False)
        s_f, L_f = tuned["choice"]
        log(f"[bench] tuned {tuned}")
    s_b, L_b = s_f, L_f
    from torchlpc_amd._ext import plan_of

    plan_f, plan_b = plan_of(x, A, s_f, L_f), plan_of(w, A, s_b, L_b, backward=True)

    stream = torch.cuda.current_stream()
    n = args.steps
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(n)]

    # the step is what LPC.apply + backward run: the forward keeps its workspace
    # (chunk transition matrices) and the backward, with the forward's strategy
    # and chunk length, reuses it
    def step(e=None):
        if e:
            e[0].record(stream)
        y, ws = ops.lpc_keep(x, A, zi, s_f, L_f)
        if e:
            e[1].record(stream)
        if ws.numel() > 0:
            out = ops.lpc_backward(w, A, zi, y, True, True, True, s_f, L_f, ws)
        else:
            out = ops.lpc_backward(w, A, zi, y, True, True, True, s_b, L_b)
        if e:
            e[2].record(stream)
        return y, out

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] timing {n} steps")
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    t_start.record(stream)
    for i in range(n):
        step(ev[i])
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    ms = total_ms / n
    fwd_ms = statistics.mean(e[0].elapsed_time(e[1]) for e in ev)
    bwd_ms = statistics.mean(e[1].elapsed_time(e[2]) for e in ev)
    step_median_ms = statistics.median(e[0].elapsed_time(e[2]) for e in ev)

    S = B * T
    bytes_fwd = (N + 2) * 4 * S
    bytes_bwd = (2 * N + 3) * 4 * S
    bytes_step = (3 * N + 5) * 4 * S
    if bwd_ms >= fwd_ms:
        dom, dom_bytes, dom_ms = "lpc_backward (fused adjoint)", bytes_bwd, bwd_ms
    else:
        dom, dom_bytes, dom_ms = "lpc (forward)", bytes_fwd, fwd_ms
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    msps = S / (ms * 1e-3) / 1e6
    log(f"[bench] step {ms:.3f} ms  fwd {fwd_ms:.3f}  bwd {bwd_ms:.3f}  -> {msps:.1f} Msamples/s")

    result = {
        "metric": "fwd+bwd Msamples/s and % HBM roofline (B=256,T=48k,N=16 fp32, ≤1e-5 rel)",
        "value": msps,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {
            "workload": f"time-varying all-pole filter fwd + fused adjoint bwd, B={B}, T={T}, N={N}, fp32",
            "B": B,
            "T": T,
            "N": N,
            "strategy_fwd": [s_f, L_f],
            "strategy_bwd": [s_b, L_b],
            "plan_fwd": list(plan_f),
            "plan_bwd": list(plan_b),
            "tuned": tuned,
        },
        "hbm_frac_step": bytes_step / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "fwd_ms": fwd_ms,
        "bwd_ms": bwd_ms,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": committed_traffic(dom, B, T, N, plan_b[0] if dom.startswith("lpc_backward") else plan_f[0]),
            "algorithmic_bytes": dom_bytes,
            "kernels": list(OP_KERNELS_BY_STRATEGY.get(plan_b[0] if dom.startswith("lpc_backward") else plan_f[0], {}).get(dom, ())),
        },
    }

    # cold steps: each one after a 1 GiB write that evicts the L2s and the MALL,
    # timed alone (the median of 10); the working set of a step (>2.6 GB) is far
    # beyond the MALL anyway, so warm and cold agree closely
    flush = torch.empty(COPY_BYTES // 4, device=dev)
    cold = []
    for _ in range(10):
        flush.fill_(1.0)
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record(stream)
        step()
        c1.record(stream)
        torch.cuda.synchronize()
        cold.append(c0.elapsed_time(c1))
    del flush
    result["step_median_ms"] = step_median_ms
    result["cold_step_median_ms"] = statistics.median(cold)

    # a device copy of COPY_BYTES, timed the same way: the bandwidth a plain
    # streaming kernel reaches on this GPU, reported beside the published peak
    src = torch.ones(COPY_BYTES // 4, device=dev)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(10):
        dst.copy_(src)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 2 * COPY_BYTES * 10 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src, dst
    result["roofline"]["measured_copy_gbs"] = copy_gbs
    result["roofline"]["frac_of_copy"] = achieved / copy_gbs
    log(f"[bench] device copy {copy_gbs:.0f} GB/s (read + write)")

    if args.calibrate_copy:
        calibrate_copy()

    if not args.no_plain:
        log("[bench] plain GPU form (forward kernel + unfused reference-style backward)")
        def plain():
            y = ops.lpc(x, A, zi, s_f, L_f)
            with torch.enable_grad():
                composite_backward(w, A, zi, y, True, True, True)
        plain()
        torch.cuda.synchronize()
        reps = 3
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            plain()
        e1.record(stream)
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1) / reps
        result["plain_gpu"] = {"ms_per_step": plain_ms, "speedup_fused": plain_ms / ms}
        log(f"[bench] plain GPU form {plain_ms:.3f} ms  (fused speedup {plain_ms / ms:.2f}x)")

    if not args.no_secondary:
        # the parity workloads, timed in the same run (same rules: inputs resident,
        # warm-up, event-timed steps on one stream); one short record each
        sec = {}
        for name in ("b16-t8192-n4", "b4-t1048576-n2", "ragged512", "ti-b1024-t16384-n32"):
            log(f"[bench] secondary workload {name}")
            r = bench_workload(name, 30, 5, tune=not args.no_tune)
            sec[name] = {
                "workload": r["config"]["workload"],
                **({"chunk_len_by_order": r["config"]["chunk_len_by_order"]} if "chunk_len_by_order" in r["config"] else {}),
                "ms_per_step": r["ms_per_step"],
                "value": r["value"],
                "unit": r["unit"],
                "hbm_frac": r["roofline"]["frac"],
                **({"rebuild_y": r["rebuild_y"]} if "rebuild_y" in r else {}),
            }
            torch.cuda.empty_cache()
        result["secondary"] = sec

    if not args.no_cpu_baseline:
        log("[bench] accuracy of the timed plan on a subset of signals")
        outs = step()
        torch.cuda.synchronize()
        sig = sorted({round(i * (B - 1) / 7) for i in range(8)})
        result["rel_err"] = accuracy_sample(x, A, zi, w, outs, sig)
        del outs
        log(f"[bench] rel_err {result['rel_err']}")
        log("[bench] reference CPU path")
        result["cpu_baseline"] = cpu_baseline(T, N)
        log(f"[bench] cpu {result['cpu_baseline']['value']:.3f} Msamples/s")

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
