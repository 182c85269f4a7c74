"""Serving runtime: continuous admission under an HBM budget, overlapping
requests on several streams; every request's output equals the same request
evaluated alone through the per-call entry point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fastfunctionsecretsharing_amd import _lib, workloads as W

    torch.cuda.set_device(0)
    templates = W.serving_templates(count=12, seed=9, lo=6, hi=13)
    return torch, _lib, W, templates


def _alone(torch, W, templates, picks):
    holder, sp = W.stream_ptr()
    outs = []
    for p in picks:
        t = templates[p]
        o = torch.empty(t.out_elems(), dtype=W.torch_dtype(t.es), device="cuda:0")
        W.eval_template_alone(t, o, sp)
        outs.append(o)
    torch.cuda.synchronize()
    del holder
    return outs


@pytest.mark.parametrize("budget,inflight,streams,mode", [(1 << 34, 16, 4, 1), (1 << 20, 4, 2, 1), (1 << 34, 1, 1, 0)])
def test_mix_equals_alone(env, budget, inflight, streams, mode):
    torch, _lib, W, templates = env
    rng = np.random.default_rng(budget % 97 + inflight)
    picks = [int(p) for p in rng.integers(0, len(templates), size=40)]
    arrivals = np.cumsum(rng.exponential(0.0005, size=len(picks)))
    arrivals[:10] = 0.0  # a burst: these are admitted together and overlap
    outs, lats, wall, st = W.run_trace(templates, picks, arrivals, budget, inflight, streams, mode)
    assert st.completed == len(picks) and st.failed == 0
    assert all(l >= 0 for l in lats)
    if inflight > 1 and budget > (1 << 30):
        assert st.maxInflightSeen > 1
    ref = _alone(torch, W, templates, picks)
    for i, (a, b) in enumerate(zip(outs, ref)):
        assert torch.equal(a, b), (i, templates[picks[i]].op, templates[picks[i]].n)


def test_mix_with_dpf_equals_alone(env):
    torch, _lib, W, _ = env
    templates = W.serving_templates(count=8, seed=21, lo=6, hi=12,
                                    ops=(_lib.SERVE_DPF, _lib.SERVE_DCF, _lib.SERVE_MIC, _lib.SERVE_SPLINE))
    rng = np.random.default_rng(3)
    picks = [int(p) for p in rng.integers(0, len(templates), size=24)]
    outs, lats, wall, st = W.run_trace(templates, picks, np.zeros(len(picks)), 1 << 34, 8, 4, 1)
    assert st.completed == len(picks) and st.failed == 0
    ref = _alone(torch, W, templates, picks)
    for i, (a, b) in enumerate(zip(outs, ref)):
        assert torch.equal(a, b), (i, templates[picks[i]].op)


def test_submit_validation(env):
    import ctypes

    torch, _lib, W, templates = env
    lib = _lib.load()
    srv = ctypes.c_void_p()
    assert lib.FastFss_hip_serverCreate(ctypes.byref(srv), 1 << 30, 4, 2, 1) == 0
    t = templates[0]
    out = torch.empty(t.out_elems(), dtype=W.torch_dtype(t.es), device="cuda:0")
    r = W.serve_request(t, out)
    r.partyId = 5
    tk = ctypes.c_uint64()
    assert lib.FastFss_hip_serverSubmit(srv, ctypes.byref(r), ctypes.byref(tk)) == -12
    r.partyId = 0
    r.keyDataSize += 1
    assert lib.FastFss_hip_serverSubmit(srv, ctypes.byref(r), ctypes.byref(tk)) == -2
    r.keyDataSize -= 1
    assert lib.FastFss_hip_serverSubmit(srv, ctypes.byref(r), ctypes.byref(tk)) == 0
    assert lib.FastFss_hip_serverWait(srv, tk.value, None) == 0
    assert lib.FastFss_hip_serverQuery(srv, tk.value) == 1
    # a released ticket is forgotten; releasing it again fails
    assert lib.FastFss_hip_serverRelease(srv, tk.value) == 0
    assert lib.FastFss_hip_serverQuery(srv, tk.value) == -1
    assert lib.FastFss_hip_serverRelease(srv, tk.value) == -1
    assert lib.FastFss_hip_serverRelease(srv, tk.value + 1000) == -1
    lib.FastFss_hip_serverDestroy(srv)


def test_concurrent_submitters(env):
    """Four host threads submit into one server at once; every request still
    equals its alone result, and each ticket is released after its wait."""
    import ctypes
    import threading

    torch, _lib, W, templates = env
    lib = _lib.load()
    rng = np.random.default_rng(17)
    picks = [[int(p) for p in rng.integers(0, len(templates), size=10)] for _ in range(4)]
    srv = ctypes.c_void_p()
    assert lib.FastFss_hip_serverCreate(ctypes.byref(srv), 1 << 34, 8, 4, 1) == 0
    outs = [[None] * 10 for _ in range(4)]
    tickets = [[0] * 10 for _ in range(4)]
    reqs = [[None] * 10 for _ in range(4)]  # kept alive until the server is done
    errors = []

    def submitter(k):
        try:
            for i, p in enumerate(picks[k]):
                t = templates[p]
                outs[k][i] = torch.empty(t.out_elems(), dtype=W.torch_dtype(t.es), device="cuda:0")
                reqs[k][i] = W.serve_request(t, outs[k][i])
                tk = ctypes.c_uint64()
                assert lib.FastFss_hip_serverSubmit(srv, ctypes.byref(reqs[k][i]), ctypes.byref(tk)) == 0
                tickets[k][i] = tk.value
        except Exception as ex:  # surfaced below on the test thread
            errors.append(ex)

    torch.cuda.synchronize()  # output allocations are complete before any request runs
    threads = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    ids = [tk for row in tickets for tk in row]
    assert len(set(ids)) == len(ids)
    for tk in ids:
        assert lib.FastFss_hip_serverWait(srv, tk, None) == 0
        assert lib.FastFss_hip_serverRelease(srv, tk) == 0
    lib.FastFss_hip_serverDestroy(srv)
    for k in range(4):
        ref = _alone(torch, W, templates, picks[k])
        for i in range(10):
            assert torch.equal(outs[k][i], ref[i]), (k, i)
