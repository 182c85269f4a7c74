"""Serving runtime: continuous admission under an HBM budget, overlapping
requests on several streams; every request's output equals the same request
evaluated alone through the per-call entry point."""
import numpy as np
import pytest

from fss_testlib import DT, OracleBackend, dcf_eval, mic_eval, spline_eval, sub_key_rows

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


def test_fused_dcf_dpf_burst(env):
    """A burst of DCF and DPF requests of mixed widths: requests admitted together share fused
    batch launches (fewer launches than requests), and each equals its alone result."""
    torch, _lib, W, _ = env
    templates = W.serving_templates(count=10, seed=33, lo=5, hi=14, widths=(8, 16, 32, 64),
                                    ops=(_lib.SERVE_DCF, _lib.SERVE_DPF))
    rng = np.random.default_rng(8)
    picks = [int(p) for p in rng.integers(0, len(templates), size=48)]
    outs, lats, wall, st = W.run_trace(templates, picks, np.zeros(len(picks)), 1 << 34, 16, 4, 1)
    assert st.completed == len(picks) and st.failed == 0
    assert st.fusedRequests >= 2 and st.launches < len(picks), (st.launches, st.fusedRequests)
    ref = _alone(torch, W, templates, picks)
    for i, (a, b) in enumerate(zip(outs, ref)):
        assert torch.equal(a, b), (i, templates[picks[i]].op, templates[picks[i]].n, templates[picks[i]].N)


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


def _np(torch, t, es):
    return t.view(torch.uint8).cpu().numpy().view(DT[es])


def test_serving_bench_trace(env):
    """The serving workload at its stated sizes: the 48 resident key sets and the 400-request
    Poisson trace (400 req/s) the benchmark serves, DCF/MIC/spline 1:1:1, 2^10..2^20 elements,
    bitWidthIn in {16, 32, 64}, under the benchmark's served configuration. Every request equals
    its alone result, and 64 elements of every key set's served output equal the oracle's."""
    torch, _lib, W, _ = env
    templates = W.serving_templates(count=48, seed=5)
    rng = np.random.default_rng(17)
    picks = [int(p) for p in rng.integers(0, len(templates), size=400)]
    arrivals = np.cumsum(rng.exponential(1.0 / 400, size=400))
    sizes = [templates[p].N for p in picks]
    assert min(sizes) >= 1 << 10 and max(sizes) <= 1 << 20 and max(sizes) > 1 << 19
    outs, lats, wall, st = W.run_trace(templates, picks, arrivals, 32 << 30, 16, 4, 1)
    assert st.completed == len(picks) and st.failed == 0
    print(f"served 400 requests in {wall:.2f} s, max in flight {st.maxInflightSeen}", flush=True)
    ref = _alone(torch, W, templates, picks)
    for i, (a, b) in enumerate(zip(outs, ref)):
        assert torch.equal(a, b), (i, templates[picks[i]].op, templates[picks[i]].n, templates[picks[i]].N)
    OR = OracleBackend()
    first = {}
    for i, p in enumerate(picks):
        first.setdefault(p, i)
    checked = set()
    for p, i in sorted(first.items()):
        t = templates[p]
        es, n, N = t.es, t.n, t.N
        idx = np.unique(np.concatenate([np.random.default_rng(p).choice(N, 62, replace=False), [0, N - 1]]))
        key = t.key.cpu().numpy()
        x = _np(torch, t.masked, es)[idx]
        seed = t.seed0.cpu().numpy().reshape(N, 16)[idx].ravel()
        got = _np(torch, outs[i], es)
        if t.op == _lib.SERVE_DCF:
            sub = sub_key_rows(key, idx, 16 * n, (n + 3) * es, N)
            want = dcf_eval(OR, x, sub, seed, 0, n, n, 1, es)
            assert np.array_equal(got[idx], want), ("dcf", p)
        elif t.op == _lib.SERVE_MIC:
            I = t.intervals
            sub = sub_key_rows(key, idx, 16 * n, (n + 3) * es, N)
            z = _np(torch, t.z, es).reshape(N, I)[idx].ravel()
            lf, rt = _np(torch, t.left, es), _np(torch, t.right, es)
            want = mic_eval(OR, x, sub, z, seed, 0, lf, rt, n, n, es)
            assert np.array_equal(got.reshape(N, I)[idx].ravel(), want), ("mic", p)
        else:
            I, d = t.intervals, t.degree
            g = I * (d + 1)
            sub = sub_key_rows(key, idx, 16 * n, (n * g + 2 + g) * es, N)
            e = _np(torch, t.z, es).reshape(N, g)[idx].ravel()
            b = _np(torch, t.beta, es).reshape(N, g)[idx].ravel()
            lf, rt = _np(torch, t.left, es), _np(torch, t.right, es)
            want = spline_eval(OR, x, sub, e, b, seed, 0, lf, rt, d, n, n, es)
            assert np.array_equal(got[idx], want), ("spline", p)
        checked.add((t.op, n))
    assert {op for op, _ in checked} == {_lib.SERVE_DCF, _lib.SERVE_MIC, _lib.SERVE_SPLINE}
    assert {n for _, n in checked} == {16, 32, 64}
