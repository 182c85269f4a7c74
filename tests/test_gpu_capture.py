"""Per-call entry points inside a captured graph.

The DCF eval runs on the work queue from 2^20 elements (or when variant 6 is
selected, as here) and the DPF eval from 2^21. Each call
then takes a 4-byte tile counter from the stream-ordered allocator, clears it
with a one-thread kernel (not a memset) and frees it on the same stream. Such
a call can be recorded into a graph. Each test captures one call, replays the
graph several times with eager work and synchronizations in between, and
checks every replay against the same call run eagerly."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fastfunctionsecretsharing_amd import _lib, workloads as W

    torch.cuda.set_device(0)
    return torch, _lib, W


def _replays_match(torch, call, out, reps=3):
    side = torch.cuda.Stream()
    holder = ctypes.c_void_p(side.cuda_stream)
    sp = ctypes.addressof(holder)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        call(sp)
    side.synchronize()
    want = out.clone()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        call(sp)
    torch.cuda.synchronize()
    for r in range(reps):
        out.zero_()
        junk = torch.arange(1 << 20, device="cuda:0").sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, want), r
        del junk
    del holder


@pytest.mark.parametrize("op", ["dcf", "dpf"])
def test_queue_eval_in_graph(env, op):
    torch, _lib, W = env
    lib = _lib.load()
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(61)
    holder, sp0 = W.stream_ptr()
    kind = _lib.BATCH_DCF_EVAL if op == "dcf" else _lib.BATCH_DPF_EVAL
    # queue form (DPF from 2^21 elements, DCF with variant 6 below), partial last tile
    g = W.make_group(kind, 32, (1 << (16 if op == "dcf" else 21)) + 77, gen, dev, sp0)
    torch.cuda.synchronize()
    out = torch.empty(g.N, dtype=g.alpha.dtype, device=dev)
    fn = lib.FastFss_cuda_dcfEval if op == "dcf" else lib.FastFss_cuda_dpfEval

    def call(sp):
        rc = fn(out.data_ptr(), 4 * g.N, g.masked.data_ptr(), 4 * g.N, g.key.data_ptr(), g.key_size,
                g.seeds[1].data_ptr(), 16 * g.N, 1, 32, 32, 1, 4, g.N, None, 0, sp)
        assert rc == 0, rc

    if op == "dcf":  # the DCF default moves to the queue only from 2^20 elements
        assert lib.FastFss_hip_setEvalVariant(6) == 0
    try:
        _replays_match(torch, call, out)
    finally:
        lib.FastFss_hip_setEvalVariant(-1)
    del holder
