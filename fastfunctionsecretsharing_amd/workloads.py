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


# ---------------------------------------------------------------------------
# serving workload: DCF / MIC / spline requests with resident keys
# ---------------------------------------------------------------------------

MIC_INTERVALS = 16
SPLINE_INTERVALS, SPLINE_DEGREE = 4, 2


@dataclass
class Template:
    op: int  # _lib.SERVE_*
    n: int
    es: int
    N: int
    masked: object
    key: object
    key_size: int
    seed0: object
    z: object = None  # MIC z / spline e (party 0's share: the whole value)
    beta: object = None
    left: object = None
    right: object = None
    intervals: int = 0
    degree: int = 0

    def out_elems(self):
        return self.N * (self.intervals if self.op == _lib.SERVE_MIC else 1)


def _sorted_intervals(rng, I, n):
    """Contiguous intervals covering [0, 2^n): left_0 = 0, right_{I-1} = 2^n - 1."""
    M = (1 << n) - 1
    cuts = sorted(set(int(v) for v in rng.integers(0, M, size=4 * I, dtype=np.uint64)))[: I - 1]
    while len(cuts) < I - 1:
        cuts.append(cuts[-1] + 1 if cuts else 0)
    right = cuts + [M]
    left = [0] + [c + 1 for c in cuts]
    return left, right


def make_template(op, n, N, rng, gen, dev, sp):
    import torch

    lib = _lib.load()
    es = 2 if n <= 16 else (4 if n <= 32 else 8)
    dt = torch_dtype(es)
    alpha = _rand_bits(gen, N, n, dev).to(dt)
    masked = _rand_bits(gen, N, n, dev).to(dt)
    seeds = torch.randint(0, 256, (2, 16 * N), dtype=torch.uint8, device=dev, generator=gen)
    ks = ctypes.c_size_t(0)

    def ep(vals):
        return torch.tensor([v if v < 2**63 else v - 2**64 for v in vals], dtype=torch.int64).to(dt).to(dev)

    if op in (_lib.SERVE_DCF, _lib.SERVE_DPF):
        dcf = op == _lib.SERVE_DCF
        size_fn = lib.FastFss_dcfGetKeyDataSize if dcf else lib.FastFss_dpfGetKeyDataSize
        assert size_fn(ctypes.addressof(ks), n, n, 1, es, N) == 0
        key = torch.empty(ks.value, dtype=torch.uint8, device=dev)
        gen_fn = lib.FastFss_cuda_dcfKeyGen if dcf else lib.FastFss_cuda_dpfKeyGen
        rc = gen_fn(key.data_ptr(), ks.value, alpha.data_ptr(), es * N, None, 0, seeds[0].data_ptr(), 16 * N,
                    seeds[1].data_ptr(), 16 * N, n, n, 1, es, N, sp)
        assert rc == 0, rc
        return Template(op, n, es, N, masked, key, ks.value, seeds[0])
    I = MIC_INTERVALS if op == _lib.SERVE_MIC else SPLINE_INTERVALS
    left, right = _sorted_intervals(rng, I, n)
    lt, rt = ep(left), ep(right)
    if op == _lib.SERVE_MIC:
        assert lib.FastFss_dcfMICGetKeyDataSize(ctypes.addressof(ks), n, n, es, N) == 0
        key = torch.empty(ks.value, dtype=torch.uint8, device=dev)
        z = torch.empty(N * I, dtype=dt, device=dev)
        rc = lib.FastFss_cuda_dcfMICKeyGen(key.data_ptr(), ks.value, z.data_ptr(), z.numel() * es, alpha.data_ptr(),
                                           es * N, seeds[0].data_ptr(), 16 * N, seeds[1].data_ptr(), 16 * N,
                                           lt.data_ptr(), es * I, rt.data_ptr(), es * I, n, n, es, N, sp)
        assert rc == 0, rc
        return Template(op, n, es, N, masked, key, ks.value, seeds[0], z=z, left=lt, right=rt, intervals=I)
    d = SPLINE_DEGREE
    g = I * (d + 1)
    coeff = _rand_bits(gen, g, n, dev).to(dt)
    assert lib.FastFss_dcfSplineGetKeyDataSize(ctypes.addressof(ks), d, I, n, n, es, N) == 0
    key = torch.empty(ks.value, dtype=torch.uint8, device=dev)
    e = torch.empty(N * g, dtype=dt, device=dev)
    beta = torch.empty(N * g, dtype=dt, device=dev)
    rc = lib.FastFss_cuda_dcfSplineKeyGen(key.data_ptr(), ks.value, e.data_ptr(), e.numel() * es, beta.data_ptr(),
                                          beta.numel() * es, alpha.data_ptr(), es * N, seeds[0].data_ptr(), 16 * N,
                                          seeds[1].data_ptr(), 16 * N, coeff.data_ptr(), g * es, d, lt.data_ptr(),
                                          es * I, rt.data_ptr(), es * I, I, n, n, es, N, sp)
    assert rc == 0, rc
    return Template(op, n, es, N, masked, key, ks.value, seeds[0], z=e, beta=beta, left=lt, right=rt,
                    intervals=I, degree=d)


def serving_templates(count=48, seed=5, lo=10, hi=20, widths=(16, 32, 64), device="cuda:0",
                      ops=(_lib.SERVE_DCF, _lib.SERVE_MIC, _lib.SERVE_SPLINE)):
    """`count` resident key sets, op round robin over `ops`, size log-uniform in
    [2^lo, 2^hi], bitWidthIn = bitWidthOut drawn from `widths`."""
    import torch

    dev = torch.device(device)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev).manual_seed(seed)
    holder, sp = stream_ptr()
    out = []
    for i in range(count):
        op = ops[i % len(ops)]
        N = int(round(math.exp(rng.uniform(math.log(2**lo), math.log(2**hi)))))
        n = int(rng.choice(widths))
        out.append(make_template(op, n, N, rng, gen, dev, sp))
    torch.cuda.synchronize()
    del holder
    return out


def serve_request(t, out):
    r = _lib.ServeRequest()
    r.op, r.partyId = t.op, 0
    r.bitWidthIn = r.bitWidthOut = t.n
    r.elementSize, r.elementNum = t.es, t.N
    r.sharedOut = out.data_ptr()
    r.maskedX = t.masked.data_ptr()
    r.key, r.keyDataSize = t.key.data_ptr(), t.key_size
    r.seed = t.seed0.data_ptr()
    if t.op in (_lib.SERVE_MIC, _lib.SERVE_SPLINE):
        r.sharedZ = t.z.data_ptr()
        r.leftEndpoints, r.rightEndpoints = t.left.data_ptr(), t.right.data_ptr()
        r.intervalNum = t.intervals
    if t.op == _lib.SERVE_SPLINE:
        r.sharedBeta, r.degree = t.beta.data_ptr(), t.degree
    return r


def template_spline_cache(t):
    """The cache buffer FastFss_cuda_dcfSplineEval takes for serving template `t`."""
    import torch

    cs = ctypes.c_size_t(0)
    assert _lib.load().FastFss_dcfSplineGetCacheDataSize(ctypes.addressof(cs), t.degree, t.intervals, t.n, t.n, t.es,
                                                         t.N) == 0
    return torch.empty(cs.value, dtype=torch.uint8, device=t.key.device)


def eval_template_alone(t, out, sp, cache=None):
    """The per-call public entry point for one request (the reference-shaped launch).  A spline
    call without `cache` allocates one and synchronizes before releasing it."""
    import torch

    lib = _lib.load()
    es, N, n = t.es, t.N, t.n
    if t.op in (_lib.SERVE_DCF, _lib.SERVE_DPF):
        fn = lib.FastFss_cuda_dcfEval if t.op == _lib.SERVE_DCF else lib.FastFss_cuda_dpfEval
        rc = fn(out.data_ptr(), es * N, t.masked.data_ptr(), es * N, t.key.data_ptr(), t.key_size,
                t.seed0.data_ptr(), 16 * N, 0, n, n, 1, es, N, None, 0, sp)
    elif t.op == _lib.SERVE_MIC:
        I = t.intervals
        rc = lib.FastFss_cuda_dcfMICEval(out.data_ptr(), es * N * I, t.masked.data_ptr(), es * N, t.key.data_ptr(),
                                         t.key_size, t.z.data_ptr(), es * N * I, t.seed0.data_ptr(), 16 * N, 0,
                                         t.left.data_ptr(), es * I, t.right.data_ptr(), es * I, n, n, es, N, None, 0,
                                         sp)
    else:
        I, d = t.intervals, t.degree
        own = cache is None
        if own:
            cache = template_spline_cache(t)
        g = I * (d + 1)
        rc = lib.FastFss_cuda_dcfSplineEval(out.data_ptr(), es * N, t.masked.data_ptr(), es * N, t.key.data_ptr(),
                                            t.key_size, t.z.data_ptr(), es * N * g, t.beta.data_ptr(), es * N * g,
                                            t.seed0.data_ptr(), 16 * N, 0, t.left.data_ptr(), es * I,
                                            t.right.data_ptr(), es * I, I, d, n, n, es, N, cache.data_ptr(),
                                            cache.numel(), sp)
        if own:
            torch.cuda.synchronize()
            del cache
    if rc != 0:
        raise RuntimeError(f"eval failed: {rc}")


def run_trace(templates, picks, arrivals, budget, max_inflight, streams, launch_mode):
    """Submit request i (template picks[i]) at time arrivals[i] seconds; return
    (outputs, latencies, wall seconds, stats)."""
    import time

    import torch

    lib = _lib.load()
    outs = [torch.empty(templates[p].out_elems(), dtype=torch_dtype(templates[p].es), device=templates[p].key.device)
            for p in picks]
    reqs = [serve_request(templates[p], o) for p, o in zip(picks, outs)]
    torch.cuda.synchronize()
    srv = ctypes.c_void_p()
    assert lib.FastFss_hip_serverCreate(ctypes.byref(srv), budget, max_inflight, streams, launch_mode) == 0
    tickets = []
    t0 = time.perf_counter()
    for r, at in zip(reqs, arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tk = ctypes.c_uint64()
        rc = lib.FastFss_hip_serverSubmit(srv, ctypes.byref(r), ctypes.byref(tk))
        assert rc == 0, rc
        tickets.append(tk.value)
    assert lib.FastFss_hip_serverDrain(srv) == 0
    wall = time.perf_counter() - t0
    lats = []
    for tk in tickets:
        lat = ctypes.c_double()
        assert lib.FastFss_hip_serverWait(srv, tk, ctypes.byref(lat)) == 0
        assert lib.FastFss_hip_serverRelease(srv, tk) == 0
        lats.append(lat.value)
    st = _lib.ServeStats()
    lib.FastFss_hip_serverStats(srv, ctypes.byref(st))
    lib.FastFss_hip_serverDestroy(srv)
    return outs, lats, wall, st


# ---------------------------------------------------------------------------
# MIC + spline batch and full-domain Grotto: DCF-MIC and a piecewise degree-2
# spline with 16 intervals at bitWidthIn = 24 (32-bit elements), bitWidthOut =
# 32, 2^22 elements; Grotto at bitWidthIn = 16 over 4096 keys, every point of
# the 2^16 domain.  Endpoints are uniform sorted contiguous intervals covering
# [0, 2^24); coefficients are uniform int32 (fixed point with 12 fractional
# bits only scales them).
# ---------------------------------------------------------------------------


@dataclass
class PiecewiseSet:
    n: int
    m: int
    es: int
    N: int
    I: int
    d: int
    alpha: object
    x: object
    masked: object
    seeds: object
    left: object
    right: object
    mic_key: object = None
    mic_ks: int = 0
    z: object = None
    spl_key: object = None
    spl_ks: int = 0
    coeff: object = None
    e: object = None
    beta: object = None


def piecewise_workload(N=1 << 22, n=24, m=32, I=16, d=2, seed=4, mic=True, spline=True, device="cuda:0"):
    """Resident MIC and spline key sets on one set of inputs (keys from the GPU keyGen entry points)."""
    import torch

    lib = _lib.load()
    dev = torch.device(device)
    es = 4
    dt = torch_dtype(es)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev).manual_seed(seed)
    holder, sp = stream_ptr()
    alpha64 = _rand_bits(gen, N, n, dev)
    x64 = _rand_bits(gen, N, n, dev)
    w = PiecewiseSet(n, m, es, N, I, d, alpha64.to(dt), x64.to(dt), ((x64 + alpha64) & ((1 << n) - 1)).to(dt),
                     torch.randint(0, 256, (2, 16 * N), dtype=torch.uint8, device=dev, generator=gen), None, None)
    left, right = _sorted_intervals(rng, I, n)
    w.left = torch.tensor(left, dtype=torch.int64).to(dt).to(dev)
    w.right = torch.tensor(right, dtype=torch.int64).to(dt).to(dev)
    ks = ctypes.c_size_t(0)
    s0, s1 = w.seeds[0].data_ptr(), w.seeds[1].data_ptr()
    if mic:
        assert lib.FastFss_dcfMICGetKeyDataSize(ctypes.addressof(ks), n, m, es, N) == 0
        w.mic_ks = ks.value
        w.mic_key = torch.empty(ks.value, dtype=torch.uint8, device=dev)
        w.z = torch.empty(N * I, dtype=dt, device=dev)
        rc = lib.FastFss_cuda_dcfMICKeyGen(w.mic_key.data_ptr(), ks.value, w.z.data_ptr(), N * I * es,
                                           w.alpha.data_ptr(), es * N, s0, 16 * N, s1, 16 * N, w.left.data_ptr(),
                                           es * I, w.right.data_ptr(), es * I, n, m, es, N, sp)
        assert rc == 0, rc
    if spline:
        g = I * (d + 1)
        w.coeff = torch.randint(-(2**31), 2**31 - 1, (g,), dtype=torch.int64, device=dev, generator=gen).to(dt)
        assert lib.FastFss_dcfSplineGetKeyDataSize(ctypes.addressof(ks), d, I, n, m, es, N) == 0
        w.spl_ks = ks.value
        w.spl_key = torch.empty(ks.value, dtype=torch.uint8, device=dev)
        w.e = torch.empty(N * g, dtype=dt, device=dev)
        w.beta = torch.empty(N * g, dtype=dt, device=dev)
        rc = lib.FastFss_cuda_dcfSplineKeyGen(w.spl_key.data_ptr(), ks.value, w.e.data_ptr(), N * g * es,
                                              w.beta.data_ptr(), N * g * es, w.alpha.data_ptr(), es * N, s0, 16 * N,
                                              s1, 16 * N, w.coeff.data_ptr(), g * es, d, w.left.data_ptr(), es * I,
                                              w.right.data_ptr(), es * I, I, n, m, es, N, sp)
        assert rc == 0, rc
    torch.cuda.synchronize()
    del holder
    return w


def mic_eval(w, party, out, z_share, sp):
    lib = _lib.load()
    N, I, es = w.N, w.I, w.es
    rc = lib.FastFss_cuda_dcfMICEval(out.data_ptr(), es * N * I, w.masked.data_ptr(), es * N, w.mic_key.data_ptr(),
                                     w.mic_ks, z_share.data_ptr(), es * N * I, w.seeds[party].data_ptr(), 16 * N,
                                     party, w.left.data_ptr(), es * I, w.right.data_ptr(), es * I, w.n, w.m, es, N,
                                     None, 0, sp)
    if rc != 0:
        raise RuntimeError(f"FastFss_cuda_dcfMICEval failed : {rc}")


def spline_cache(w):
    import torch

    lib = _lib.load()
    cs = ctypes.c_size_t(0)
    assert lib.FastFss_dcfSplineGetCacheDataSize(ctypes.addressof(cs), w.d, w.I, w.n, w.m, w.es, w.N) == 0
    return torch.empty(cs.value, dtype=torch.uint8, device=w.spl_key.device)


def spline_eval(w, party, out, e_share, beta_share, cache, sp):
    lib = _lib.load()
    N, I, es, d = w.N, w.I, w.es, w.d
    g = I * (d + 1)
    rc = lib.FastFss_cuda_dcfSplineEval(out.data_ptr(), es * N, w.masked.data_ptr(), es * N, w.spl_key.data_ptr(),
                                        w.spl_ks, e_share.data_ptr(), es * N * g, beta_share.data_ptr(), es * N * g,
                                        w.seeds[party].data_ptr(), 16 * N, party, w.left.data_ptr(), es * I,
                                        w.right.data_ptr(), es * I, I, d, w.n, w.m, es, N, cache.data_ptr(),
                                        cache.numel(), sp)
    if rc != 0:
        raise RuntimeError(f"FastFss_cuda_dcfSplineEval failed : {rc}")


@dataclass
class GrottoSet:
    n: int
    es: int
    N: int
    alpha: object
    seeds: object
    key: object
    ks: int


def grotto_keys(N=4096, n=16, seed=6, device="cuda:0"):
    import torch

    lib = _lib.load()
    dev = torch.device(device)
    es = 2 if n <= 16 else (4 if n <= 32 else 8)
    gen = torch.Generator(device=dev).manual_seed(seed)
    holder, sp = stream_ptr()
    alpha = _rand_bits(gen, N, n, dev).to(torch_dtype(es))
    seeds = torch.randint(0, 256, (2, 16 * N), dtype=torch.uint8, device=dev, generator=gen)
    ks = ctypes.c_size_t(0)
    assert lib.FastFss_grottoGetKeyDataSize(ctypes.addressof(ks), n, es, N) == 0
    key = torch.empty(ks.value, dtype=torch.uint8, device=dev)
    rc = lib.FastFss_cuda_grottoKeyGen(key.data_ptr(), ks.value, alpha.data_ptr(), es * N, seeds[0].data_ptr(),
                                       16 * N, seeds[1].data_ptr(), 16 * N, n, es, N, sp)
    assert rc == 0, rc
    torch.cuda.synchronize()
    del holder
    return GrottoSet(n, es, N, alpha, seeds, key, ks.value)


def grotto_full(gs, party, mode, packed, out, sp):
    lib = _lib.load()
    rc = lib.FastFss_hip_grottoEvalFullDomain(out.data_ptr(), out.numel() * out.element_size(), gs.key.data_ptr(),
                                              gs.ks, gs.seeds[party].data_ptr(), 16 * gs.N, party, mode, packed,
                                              gs.n, gs.es, gs.N, sp)
    if rc != 0:
        raise RuntimeError(f"FastFss_hip_grottoEvalFullDomain failed : {rc}")
