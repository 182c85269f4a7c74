"""DCF-MIC and DCF-spline at full workload size on the GPU: 16 contiguous
intervals, bitWidthIn = 24, bitWidthOut = 32, 2^22 elements.

* MIC: the two parties' shares reconstruct to [x in [p_i, q_i]] (signed
  n-bit interval test of tests/mic/mic_eval_test_case.h:84-106) for every
  element and interval, and a sample of elements matches the oracle bit for
  bit (keys, z, both shares).
* Spline: a sample of elements matches the oracle bit for bit (keys, e, beta,
  both parties' shares on random additive splits of e and beta).
"""
import numpy as np
import pytest

from fss_testlib import (DT, OracleBackend, mic_eval, mic_keygen, signed_nbit, spline_eval, spline_keygen)

pytestmark = pytest.mark.gpu

OR = OracleBackend()


@pytest.fixture(scope="module")
def work():
    import torch

    from fastfunctionsecretsharing_amd import workloads as W

    torch.cuda.set_device(0)
    return torch, W, W.piecewise_workload()


def _h(torch, t):
    torch.cuda.synchronize()
    return t.view(torch.uint8).cpu().numpy().view(DT[t.element_size()])


SAMPLE = np.sort(np.random.default_rng(3).choice(1 << 22, 1024, replace=False))


def _rows(torch, t, N, offset=0, length=None):
    """Rows SAMPLE of the (N, -1) matrix stored in t's bytes [offset, offset + length), gathered on the device."""
    b = t.view(torch.uint8)
    b = b[offset:offset + length] if length is not None else b[offset:]
    idx = torch.from_numpy(SAMPLE).to(b.device)
    r = b.view(N, -1)[idx].reshape(-1)
    return r.cpu().numpy().view(DT[t.element_size()] if length is None and offset == 0 else np.uint8)


def _dcf_sub_key(torch, key, n, g, es, N):
    """Elements SAMPLE of a DCF-layout key (src/impl/dcf.h:139-167)."""
    sc = _rows(torch, key, N, 0, N * 16 * n)
    tl = _rows(torch, key, N, N * 16 * n, N * es * (n * g + 2 + g))
    return np.concatenate([sc, tl])


def test_mic_full_size(work):
    torch, W, w = work
    N, I, n = w.N, w.I, w.n
    holder, sp = W.stream_ptr()
    gen = torch.Generator(device="cuda:0").manual_seed(8)
    z0 = torch.randint(-(2**31), 2**31 - 1, (N * I,), dtype=torch.int64, device="cuda:0", generator=gen).to(
        torch.int32)
    z1 = w.z - z0
    y = []
    for party, zs in ((0, z0), (1, z1)):
        out = torch.empty(N * I, dtype=torch.int32, device="cuda:0")
        W.mic_eval(w, party, out, zs, sp)
        y.append(out)
    rec = (y[0].to(torch.int64) + y[1].to(torch.int64)) & 0xFFFFFFFF
    half = 1 << (n - 1)
    sx = w.x.to(torch.int64) & ((1 << n) - 1)
    sx = torch.where(sx >= half, sx - (1 << n), sx)
    L = [signed_nbit(int(v) & 0xFFFFFF, n) for v in _h(torch, w.left)]
    R = [signed_nbit(int(v) & 0xFFFFFF, n) for v in _h(torch, w.right)]
    rec = rec.reshape(N, I)
    for i in range(I):
        if L[i] > R[i]:
            continue  # wraps in signed order; the reference's test grid never forms such intervals
        want = ((sx >= L[i]) & (sx <= R[i])).to(torch.int64)
        assert torch.equal(rec[:, i], want), i
    # bit-exact against the oracle on a sample
    es = w.es
    sub = _dcf_sub_key(torch, w.mic_key, n, 1, es, N)
    a_h = _rows(torch, w.alpha, N)
    s0, s1 = _rows(torch, w.seeds[0], N), _rows(torch, w.seeds[1], N)
    lf, rt = _h(torch, w.left), _h(torch, w.right)
    ko, zo = mic_keygen(OR, a_h, s0, s1, lf, rt, n, w.m, es)
    assert np.array_equal(ko, sub)
    assert np.array_equal(zo, _rows(torch, w.z, N))
    m_h = _rows(torch, w.masked, N)
    for party, zs, sd in ((0, z0, s0), (1, z1, s1)):
        want = mic_eval(OR, m_h, ko, _rows(torch, zs, N), sd, party, lf, rt, n, w.m, es)
        assert np.array_equal(_rows(torch, y[party], N), want), party
    del holder


def test_spline_full_size(work):
    torch, W, w = work
    N, I, n, d, es = w.N, w.I, w.n, w.d, w.es
    g = I * (d + 1)
    holder, sp = W.stream_ptr()
    gen = torch.Generator(device="cuda:0").manual_seed(9)
    e0 = torch.randint(-(2**31), 2**31 - 1, (N * g,), dtype=torch.int64, device="cuda:0", generator=gen).to(
        torch.int32)
    b0 = torch.randint(-(2**31), 2**31 - 1, (N * g,), dtype=torch.int64, device="cuda:0", generator=gen).to(
        torch.int32)
    cache = W.spline_cache(w)
    y = []
    for party, es_, bs_ in ((0, e0, b0), (1, w.e - e0, w.beta - b0)):
        out = torch.empty(N, dtype=torch.int32, device="cuda:0")
        W.spline_eval(w, party, out, es_, bs_, cache, sp)
        y.append((out, es_, bs_))
    del cache
    sub = _dcf_sub_key(torch, w.spl_key, n, g, es, N)
    a_h = _rows(torch, w.alpha, N)
    s0, s1 = _rows(torch, w.seeds[0], N), _rows(torch, w.seeds[1], N)
    lf, rt, cf = _h(torch, w.left), _h(torch, w.right), _h(torch, w.coeff)
    ko, eo, bo = spline_keygen(OR, a_h, s0, s1, cf, d, lf, rt, n, w.m, es)
    assert np.array_equal(ko, sub)
    assert np.array_equal(eo, _rows(torch, w.e, N)) and np.array_equal(bo, _rows(torch, w.beta, N))
    m_h = _rows(torch, w.masked, N)
    for party, sd in ((0, s0), (1, s1)):
        out, es_, bs_ = y[party]
        want = spline_eval(OR, m_h, ko, _rows(torch, es_, N), _rows(torch, bs_, N), sd, party, lf, rt, d, n, w.m, es)
        assert np.array_equal(_rows(torch, out, N), want), party
    del holder
