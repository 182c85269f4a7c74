"""Synthetic workloads built on the device library (keys generated on the GPU).

``ragged_batch`` builds the mixed-width DCF/DPF key groups of the ragged
benchmark configuration: group sizes log-uniform in [2^lo, 2^hi] elements,
bitWidthIn drawn from {8, 16, 24, 32, 64} with bitWidthOut = bitWidthIn
(24-bit values stored in 32-bit elements), DCF:DPF = 70:30, inputs and seeds
uniform from a seeded generator.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

STORAGE = {8: 1, 16: 2, 24: 4, 32: 4, 64: 8}


def torch_dtype(es):
    import torch

    return {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[es]


def stream_ptr(stream=None):
    """Address of a c_void_p holding the hipStream_t of `stream` (kept alive by the caller)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    holder = ctypes.c_void_p(s.cuda_stream)
    return holder, ctypes.addressof(holder)


@dataclass
class Group:
    op: int  # _lib.BATCH_DCF_EVAL / BATCH_DPF_EVAL
    n: int
    es: int
    N: int
    alpha: object
    masked: object
    seeds: object  # (2, 16N) uint8
    key: object
    key_size: int
    out: list = field(default_factory=list)  # per party output tensors


def _rand_bits(gen, N, bits, dev):
    import torch

    v = torch.randint(0, 2**62, (N,), dtype=torch.int64, device=dev, generator=gen)
    v ^= torch.randint(0, 4, (N,), dtype=torch.int64, device=dev, generator=gen) << 62
    if bits < 64:
        v &= (1 << bits) - 1
    return v


def make_group(op, n, N, gen, dev, sp):
    """Keys for one group: alpha, masked x = x + alpha, seeds, key via keyGen on the GPU."""
    import torch

    lib = _lib.load()
    es = STORAGE[n] if n in STORAGE else (1 if n <= 8 else 2 if n <= 16 else 4 if n <= 32 else 8)
    dt = torch_dtype(es)
    alpha = _rand_bits(gen, N, n, dev)
    x = _rand_bits(gen, N, n, dev)
    masked = x + alpha
    if n < 64:
        masked &= (1 << n) - 1
    a_t, m_t = alpha.to(dt), masked.to(dt)
    seeds = torch.randint(0, 256, (2, 16 * N), dtype=torch.uint8, device=dev, generator=gen)
    ks = ctypes.c_size_t(0)
    size_fn = lib.FastFss_dcfGetKeyDataSize if op == _lib.BATCH_DCF_EVAL else lib.FastFss_dpfGetKeyDataSize
    assert size_fn(ctypes.addressof(ks), n, n, 1, es, N) == 0
    key = torch.empty(ks.value, dtype=torch.uint8, device=dev)
    gen_fn = lib.FastFss_cuda_dcfKeyGen if op == _lib.BATCH_DCF_EVAL else lib.FastFss_cuda_dpfKeyGen
    rc = gen_fn(key.data_ptr(), ks.value, a_t.data_ptr(), es * N, None, 0, seeds[0].data_ptr(), 16 * N,
                seeds[1].data_ptr(), 16 * N, n, n, 1, es, N, sp)
    assert rc == 0, rc
    return Group(op, n, es, N, a_t, m_t, seeds, key, ks.value)


def ragged_batch(groups=1024, seed=1, lo=6, hi=18, widths=(8, 16, 24, 32, 64), dcf_frac=0.7, device="cuda:0"):
    import torch

    dev = torch.device(device)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev).manual_seed(seed)
    holder, sp = stream_ptr()
    out = []
    for _ in range(groups):
        N = int(round(math.exp(rng.uniform(math.log(2**lo), math.log(2**hi)))))
        n = int(rng.choice(widths))
        op = _lib.BATCH_DCF_EVAL if rng.random() < dcf_frac else _lib.BATCH_DPF_EVAL
        out.append(make_group(op, n, N, gen, dev, sp))
    torch.cuda.synchronize()
    del holder
    return out


def batch_descriptors(groups, party, outs, tile=0):
    """FastFssBatchGroup array for `groups` evaluated by `party` into `outs`."""
    arr = (_lib.BatchGroup * len(groups))()
    for i, g in enumerate(groups):
        d = arr[i]
        d.op, d.partyId = g.op, party
        d.bitWidthIn = d.bitWidthOut = g.n
        d.elementSize, d.elementNum = g.es, g.N
        d.sharedOut = outs[i].data_ptr()
        d.maskedX = g.masked.data_ptr()
        d.key = g.key.data_ptr()
        d.seed = g.seeds[party].data_ptr()
        d.keyDataSize = g.key_size
        d.tileElems = tile
    return arr


def eval_plain(groups, party, outs, sp):
    """The per-call form: one FastFss_cuda_{dcf,dpf}Eval per group on one stream."""
    lib = _lib.load()
    for g, o in zip(groups, outs):
        fn = lib.FastFss_cuda_dcfEval if g.op == _lib.BATCH_DCF_EVAL else lib.FastFss_cuda_dpfEval
        rc = fn(o.data_ptr(), g.es * g.N, g.masked.data_ptr(), g.es * g.N, g.key.data_ptr(), g.key_size,
                g.seeds[party].data_ptr(), 16 * g.N, party, g.n, g.n, 1, g.es, g.N, None, 0, sp)
        if rc != 0:
            raise RuntimeError(f"eval failed: {rc}")


def alloc_outputs(groups):
    import torch

    return [torch.empty(g.N, dtype=torch_dtype(g.es), device=g.key.device) for g in groups]


def reconstruct_ok(groups, outs0, outs1):
    """share0 + share1 == [x < alpha] (DCF) / [x == alpha] (DPF), mod 2^n, per group."""
    import torch

    for g, a, b in zip(groups, outs0, outs1):
        M = (1 << g.n) - 1 if g.n < 64 else -1
        s = (a.to(torch.int64) + b.to(torch.int64)) & M
        am = g.alpha.to(torch.int64) & ((1 << (8 * g.es)) - 1 if g.es < 8 else -1)
        xm = g.masked.to(torch.int64) & ((1 << (8 * g.es)) - 1 if g.es < 8 else -1)
        if g.es == 8:  # unsigned compare of 64-bit patterns
            am, xm = am ^ (-(2**63)), xm ^ (-(2**63))
        want = (xm < am) if g.op == _lib.BATCH_DCF_EVAL else (xm == am)
        if not torch.equal(s, want.to(torch.int64) & M):
            return False
    return True
