"""Autotuner for the MI355X launch shapes, measured on the device.

Two knobs are tuned by timing the real kernels:

* ``tune_batch_tiles``: queue tile size (elements per tile, multiple of 64)
  of the ragged-batch kernel per (operation, element size).  It runs
  coordinate descent on the batch the caller is about to run. One plan is
  built per candidate table and replayed from its hipGraph, and a candidate
  is kept only when it beats the incumbent by more than ``min_gain``.
* ``tune_eval_variants``: the per-call DCF eval variant (correction-word
  levels per burst x elements per lane) per element size, timed on one
  group of each width.

Results are applied through the library's setters
(``FastFss_hip_batchSetDefaultTile``, ``FastFss_hip_setEvalVariantForElementSize``)
and can be saved to / loaded from a JSON file keyed by the device name and CU
count, so a deployment tunes once per machine type.
"""
from __future__ import annotations

import ctypes
import json
import time
from pathlib import Path

from . import _lib
from . import workloads as W

TILE_CANDIDATES = (64, 128, 256, 512, 1024, 2048)


def _device_id():
    import torch

    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|cu{p.multi_processor_count}"


def _time_plan(lib, desc, n_groups, sp, stream, reps):
    import torch

    plan = ctypes.c_void_p()
    if lib.FastFss_hip_batchPlanCreate(ctypes.byref(plan), desc, n_groups) != 0:
        raise RuntimeError("FastFss_hip_batchPlanCreate failed")
    try:
        if lib.FastFss_hip_batchPlanCapture(plan, sp) != 0:
            raise RuntimeError("FastFss_hip_batchPlanCapture failed")
        lib.FastFss_hip_batchPlanReplay(plan, sp)  # warm
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            lib.FastFss_hip_batchPlanReplay(plan, sp)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps
    finally:
        lib.FastFss_hip_batchPlanDestroy(plan)


def tune_batch_tiles(groups, party=0, candidates=TILE_CANDIDATES, reps=3, sweeps=1, min_gain=0.01, apply=True):
    """Pick the queue tile per (op, element size) for `groups` (workloads.Group list).

    Returns ({(op, es): tile}, best device ms per batch, seconds spent)."""
    import torch

    lib = _lib.load()
    t0 = time.perf_counter()
    # graph capture needs a stream other than the legacy default stream
    caller = torch.cuda.current_stream()
    stream = torch.cuda.Stream()
    stream.wait_stream(caller)
    holder, sp = W.stream_ptr(stream)
    outs = W.alloc_outputs(groups)
    desc = W.batch_descriptors(groups, party, outs)
    keys = sorted({(int(d.op), int(d.elementSize)) for d in desc})
    table = {k: int(lib.FastFss_hip_batchGetDefaultTile(k[0], k[1])) for k in keys}

    def run(tab):
        for d in desc:
            d.tileElems = tab[(int(d.op), int(d.elementSize))]
        return _time_plan(lib, desc, len(groups), sp, stream, reps)

    best = run(table)
    for _ in range(sweeps):
        for k in keys:
            for c in candidates:
                if c == table[k]:
                    continue
                trial = dict(table)
                trial[k] = c
                t = run(trial)
                if t < best * (1.0 - min_gain):
                    best, table = t, trial
    caller.wait_stream(stream)
    if apply:
        for (op, es), tile in table.items():
            lib.FastFss_hip_batchSetDefaultTile(op, es, tile)
    del holder
    return table, best, time.perf_counter() - t0


def tune_eval_variants(widths=(8, 16, 32, 64), log2n=20, reps=5, apply=True):
    """Pick the per-call DCF eval variant per element size.  Returns ({es: variant}, {es: ms})."""
    import torch

    lib = _lib.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    stream = torch.cuda.current_stream()
    holder, sp = W.stream_ptr(stream)
    gen = torch.Generator(device=dev).manual_seed(99)
    N = 1 << log2n
    best, best_ms = {}, {}
    for n in widths:
        g = W.make_group(_lib.BATCH_DCF_EVAL, n, N, gen, dev, sp)
        out = torch.empty(N, dtype=W.torch_dtype(g.es), device=dev)

        def step():
            rc = lib.FastFss_cuda_dcfEval(out.data_ptr(), g.es * N, g.masked.data_ptr(), g.es * N,
                                          g.key.data_ptr(), g.key_size, g.seeds[0].data_ptr(), 16 * N, 0, n, n, 1,
                                          g.es, N, None, 0, sp)
            if rc != 0:
                raise RuntimeError(f"FastFss_cuda_dcfEval failed : {rc}")

        for v in range(lib.FastFss_hip_getEvalVariantCount()):
            lib.FastFss_hip_setEvalVariantForElementSize(g.es, v)
            step()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                step()
            e1.record(stream)
            stream.synchronize()
            ms = e0.elapsed_time(e1) / reps
            key = g.es
            # widths sharing an element size add up
            best_ms.setdefault(key, {}).setdefault(v, 0.0)
            best_ms[key][v] += ms
        del g, out
    for es, per_v in best_ms.items():
        best[es] = min(per_v, key=per_v.get)
    if apply:
        for es, v in best.items():
            lib.FastFss_hip_setEvalVariantForElementSize(es, v)
    else:
        for es in best:
            lib.FastFss_hip_setEvalVariantForElementSize(es, -1)
    del holder
    return best, {es: per_v[best[es]] for es, per_v in best_ms.items()}


def save(path, tiles=None, variants=None):
    """Write tuned values for this device to `path` (merged with what is there)."""
    p = Path(path)
    data = json.loads(p.read_text()) if p.exists() else {}
    entry = data.setdefault(_device_id(), {})
    if tiles:
        entry["tiles"] = {f"{op}:{es}": t for (op, es), t in tiles.items()}
    if variants:
        entry["variants"] = {str(es): v for es, v in variants.items()}
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(data, indent=1, sort_keys=True))


def load_and_apply(path):
    """Apply the values stored for this device; returns the entry applied ({} if none)."""
    p = Path(path)
    if not p.exists():
        return {}
    entry = json.loads(p.read_text()).get(_device_id(), {})
    lib = _lib.load()
    for k, t in entry.get("tiles", {}).items():
        op, es = (int(v) for v in k.split(":"))
        lib.FastFss_hip_batchSetDefaultTile(op, es, int(t))
    for es, v in entry.get("variants", {}).items():
        lib.FastFss_hip_setEvalVariantForElementSize(int(es), int(v))
    return entry
