"""Ragged batch (one persistent launch over mixed-width DCF/DPF groups).

Each group's shares from the batch must equal the per-call entry point's
shares bit for bit, and some groups are checked against the oracle directly.
The captured hipGraph is replayed several times with eager work in between;
every replay must match the eager result.
"""
import ctypes

import numpy as np
import pytest

from fss_testlib import DT, OracleBackend, dcf_eval, dpf_eval, sub_key_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch

    from fastfunctionsecretsharing_amd import _lib, workloads as W

    torch.cuda.set_device(0)
    groups = W.ragged_batch(groups=48, seed=3, lo=0, hi=12)
    return torch, _lib, W, groups


def _run_batch(torch, lib, W, groups, party, sp, tile=0):
    outs = W.alloc_outputs(groups)
    for o in outs:
        o.fill_(0x55)
    desc = W.batch_descriptors(groups, party, outs, tile)
    assert lib.FastFss_hip_batchEval(desc, len(groups), sp) == 0
    torch.cuda.synchronize()
    return outs


def test_batch_equals_per_call(setup):
    torch, _lib, W, groups = setup
    lib = _lib.load()
    holder, sp = W.stream_ptr()
    for party in (0, 1):
        ref = W.alloc_outputs(groups)
        W.eval_plain(groups, party, ref, sp)
        for tile in (0, 64, 256):
            outs = _run_batch(torch, lib, W, groups, party, sp, tile)
            for i, (a, b) in enumerate(zip(outs, ref)):
                assert torch.equal(a, b), (party, tile, i, groups[i].op, groups[i].n, groups[i].N)
    del holder


def test_batch_reconstructs_and_matches_oracle(setup):
    torch, _lib, W, groups = setup
    lib = _lib.load()
    holder, sp = W.stream_ptr()
    o0 = _run_batch(torch, lib, W, groups, 0, sp)
    o1 = _run_batch(torch, lib, W, groups, 1, sp)
    assert W.reconstruct_ok(groups, o0, o1)
    OR = OracleBackend()
    for i, g in enumerate(groups[:12]):
        es, n = g.es, g.n
        dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[es]
        x = g.masked.view(torch.uint8).cpu().numpy().view(dt)
        key = g.key.cpu().numpy()
        for party, out in ((0, o0[i]), (1, o1[i])):
            seed = g.seeds[party].cpu().numpy()
            fn = dcf_eval if g.op == _lib.BATCH_DCF_EVAL else dpf_eval
            want = fn(OR, x, key, seed, party, n, n, 1, es)
            assert np.array_equal(out.view(torch.uint8).cpu().numpy().view(dt), want), (i, party)
    del holder


def test_batch_graph_replay(setup):
    torch, _lib, W, groups = setup
    lib = _lib.load()
    stream = torch.cuda.Stream()
    holder = ctypes.c_void_p(stream.cuda_stream)
    sp = ctypes.addressof(holder)
    outs = W.alloc_outputs(groups)
    desc = W.batch_descriptors(groups, 1, outs)
    plan = ctypes.c_void_p()
    assert lib.FastFss_hip_batchPlanCreate(ctypes.byref(plan), desc, len(groups)) == 0
    assert lib.FastFss_hip_batchPlanTileCount(plan) > 0
    with torch.cuda.stream(stream):
        assert lib.FastFss_hip_batchPlanLaunch(plan, sp) == 0
    stream.synchronize()
    eager = [o.clone() for o in outs]
    assert lib.FastFss_hip_batchPlanCapture(plan, sp) == 0
    for rep in range(4):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        assert lib.FastFss_hip_batchPlanReplay(plan, sp) == 0
        stream.synchronize()
        for a, b in zip(outs, eager):
            assert torch.equal(a, b), rep
        _ = torch.randn(1000, device="cuda:0").sum().item()  # eager work between replays
    lib.FastFss_hip_batchPlanDestroy(plan)
    del holder


def test_batch_validation(setup):
    torch, _lib, W, groups = setup
    lib = _lib.load()
    holder, sp = W.stream_ptr()
    outs = W.alloc_outputs(groups[:2])
    desc = W.batch_descriptors(groups[:2], 0, outs)
    desc[1].partyId = 3
    assert lib.FastFss_hip_batchEval(desc, 2, sp) == -12
    desc[1].partyId = 0
    desc[1].keyDataSize += 1
    assert lib.FastFss_hip_batchEval(desc, 2, sp) == -2
    desc[1].keyDataSize -= 1
    desc[0].elementSize = 3
    assert lib.FastFss_hip_batchEval(desc, 2, sp) == -11
    assert lib.FastFss_hip_batchEval(desc, 0, sp) == 0
    del holder


def test_ragged_batch_1024_groups(setup):
    """The ragged workload at its stated size: the benchmark's 1024 groups (log-uniform 2^6..2^18
    elements, bitWidthIn = bitWidthOut in {8, 16, 24, 32, 64}, DCF:DPF 70:30, 31M elements) in one
    captured plan, replayed for each party. Every group equals the per-call entry point, every
    group reconstructs, and 128 elements of one group per (op, width) equal the oracle's shares."""
    torch, _lib, W, _ = setup
    lib = _lib.load()
    groups = W.ragged_batch(groups=1024, seed=11)
    assert len(groups) == 1024 and sum(g.N for g in groups) > 30_000_000
    stream = torch.cuda.Stream()
    holder = ctypes.c_void_p(stream.cuda_stream)
    sp = ctypes.addressof(holder)
    outs = {}
    for party in (0, 1):
        o = W.alloc_outputs(groups)
        plan = ctypes.c_void_p()
        assert lib.FastFss_hip_batchPlanCreate(ctypes.byref(plan), W.batch_descriptors(groups, party, o),
                                               len(groups)) == 0
        assert lib.FastFss_hip_batchPlanCapture(plan, sp) == 0
        assert lib.FastFss_hip_batchPlanReplay(plan, sp) == 0
        stream.synchronize()
        lib.FastFss_hip_batchPlanDestroy(plan)
        ref = W.alloc_outputs(groups)
        with torch.cuda.stream(stream):
            W.eval_plain(groups, party, ref, sp)
        stream.synchronize()
        for i, (a, b) in enumerate(zip(o, ref)):
            assert torch.equal(a, b), (party, i, groups[i].op, groups[i].n, groups[i].N)
        del ref
        outs[party] = o
        print(f"party {party}: 1024 batched groups equal the per-call launches", flush=True)
    assert W.reconstruct_ok(groups, outs[0], outs[1])
    OR = OracleBackend()
    seen = set()
    for i, g in enumerate(groups):
        if (g.op, g.n) in seen or g.N < 256:
            continue
        seen.add((g.op, g.n))
        es, n, N = g.es, g.n, g.N
        idx = np.unique(np.concatenate([np.random.default_rng(i).choice(N, 126, replace=False), [0, N - 1]]))
        sub = sub_key_rows(g.key.cpu().numpy(), idx, 16 * n, (n + 3) * es if g.op == _lib.BATCH_DCF_EVAL else 3 * es, N)
        x = g.masked.view(torch.uint8).cpu().numpy().view(DT[es])[idx]
        fn = dcf_eval if g.op == _lib.BATCH_DCF_EVAL else dpf_eval
        for party in (0, 1):
            seed = g.seeds[party].cpu().numpy().reshape(N, 16)[idx].ravel()
            got = outs[party][i].view(torch.uint8).cpu().numpy().view(DT[es])[idx]
            assert np.array_equal(got, fn(OR, x, sub, seed, party, n, n, 1, es)), (i, party, g.op, n)
    assert len(seen) == 10, seen
    del holder
