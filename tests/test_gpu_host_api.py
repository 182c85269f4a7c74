"""The host-buffer entry points (FastFss_cpu_*, include/FastFss/cpu/*.h) on the GPU: numpy buffers
in host memory go through the same argument lists as the oracle's, and keys, shares and error codes
equal the oracle's, for every primitive; the PRNG's stream and counter too.  pyFastFss ops on CPU
tensors (which take this path) equal the same ops on GPU tensors."""
import ctypes

import numpy as np
import pytest

from fss_testlib import (DT, HostBackend, OracleBackend, dcf_eval, dcf_keygen, dpf_eval, dpf_eval_all,
                         dpf_eval_multi, dpf_keygen, grotto_eq_eval, grotto_eval, grotto_keygen, grotto_lut_eval,
                         grotto_mic_eval, le_block, mask, mic_eval, mic_keygen, oracle, ottt_eval, ottt_keygen,
                         rand_elems, rand_seeds, spline_eval, spline_keygen)

pytestmark = pytest.mark.gpu

OR = OracleBackend()


@pytest.fixture(scope="module")
def host():
    return HostBackend()


def _masked(x, alpha, n, es):
    return ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])


@pytest.mark.parametrize("n,m,g,es,N", [(16, 16, 1, 8, 1 << 16), (32, 32, 1, 4, 5000), (24, 24, 3, 4, 300),
                                        (8, 8, 1, 1, 77)])
def test_dcf(host, n, m, g, es, N):
    rng = np.random.default_rng(n + g + N)
    alpha, x = rand_elems(rng, N, n, es), rand_elems(rng, N, n, es)
    beta = None if g == 1 else rand_elems(rng, N * g, m, es)
    masked = _masked(x, alpha, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kh = dcf_keygen(host, alpha, beta, s0, s1, n, m, g, es)
    ko = dcf_keygen(OR, alpha, beta, s0, s1, n, m, g, es)
    assert np.array_equal(kh, ko)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dcf_eval(host, masked, kh, sd, party, n, m, g, es),
                              dcf_eval(OR, masked, ko, sd, party, n, m, g, es)), party


def test_dpf_family(host):
    rng = np.random.default_rng(5)
    N, n, m, es = 400, 12, 16, 2
    alpha, x = rand_elems(rng, N, n, es), rand_elems(rng, N, n, es)
    x[::3] = alpha[::3]
    beta = rand_elems(rng, N, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kh = dpf_keygen(host, alpha, beta, s0, s1, n, m, 1, es)
    ko = dpf_keygen(OR, alpha, beta, s0, s1, n, m, 1, es)
    assert np.array_equal(kh, ko)
    pts = rand_elems(rng, 5, n, es)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dpf_eval(host, x, kh, sd, party, n, m, 1, es), dpf_eval(OR, x, ko, sd, party, n, m, 1, es))
        assert np.array_equal(dpf_eval_multi(host, x, kh, sd, party, pts, n, m, 1, es),
                              dpf_eval_multi(OR, x, ko, sd, party, pts, n, m, 1, es))
    sub = slice(0, 6)
    kh6 = dpf_keygen(host, alpha[sub], beta[sub], s0[:96], s1[:96], n, m, 1, es)
    assert np.array_equal(dpf_eval_all(host, x[sub], kh6, s0[:96], 0, n, m, 1, es),
                          dpf_eval_all(OR, x[sub], kh6, s0[:96], 0, n, m, 1, es))


def test_grotto_family(host):
    rng = np.random.default_rng(6)
    N, n, es = 500, 16, 2
    alpha, x = rand_elems(rng, N, n, es), rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kh = grotto_keygen(host, alpha, s0, s1, n, es)
    ko = grotto_keygen(OR, alpha, s0, s1, n, es)
    assert np.array_equal(kh, ko)
    left = np.array([0, 1000, 30000], dtype=DT[es])
    right = np.array([999, 29999, 65535], dtype=DT[es])
    lut = rand_elems(rng, 6, 16, es)
    for party, sd in ((0, s0), (1, s1)):
        for eb in (False, True):
            assert np.array_equal(grotto_eval(host, x, kh, sd, party, n, es, eb), grotto_eval(OR, x, ko, sd, party, n, es, eb))
        assert np.array_equal(grotto_eq_eval(host, x, kh, sd, party, n, es), grotto_eq_eval(OR, x, ko, sd, party, n, es))
        assert np.array_equal(grotto_mic_eval(host, x, kh, sd, party, left, right, n, es),
                              grotto_mic_eval(OR, x, ko, sd, party, left, right, n, es))
        eh, th = grotto_lut_eval(host, x, kh, sd, party, left, right, lut, n, 16, es)
        eo, to = grotto_lut_eval(OR, x, ko, sd, party, left, right, lut, n, 16, es)
        assert np.array_equal(eh, eo) and np.array_equal(th, to)


def test_mic_and_spline(host):
    rng = np.random.default_rng(7)
    N, n, m, es, I, d = 300, 24, 32, 4, 5, 2
    alpha, x = rand_elems(rng, N, n, es), rand_elems(rng, N, n, es)
    masked = _masked(x, alpha, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    cuts = sorted(set(int(v) for v in rand_elems(rng, 4 * I, n, es)))[:I - 1]
    right = np.array(cuts + [mask(n)], dtype=DT[es])
    left = np.array([0] + [c + 1 for c in cuts], dtype=DT[es])
    kh, zh = mic_keygen(host, alpha, s0, s1, left, right, n, m, es)
    ko, zo = mic_keygen(OR, alpha, s0, s1, left, right, n, m, es)
    assert np.array_equal(kh, ko) and np.array_equal(zh, zo)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(mic_eval(host, masked, kh, zh, sd, party, left, right, n, m, es),
                              mic_eval(OR, masked, ko, zo, sd, party, left, right, n, m, es))
    coeff = rand_elems(rng, I * (d + 1), m, es)
    kh, eh, bh = spline_keygen(host, alpha, s0, s1, coeff, d, left, right, n, m, es)
    ko, eo, bo = spline_keygen(OR, alpha, s0, s1, coeff, d, left, right, n, m, es)
    assert np.array_equal(kh, ko) and np.array_equal(eh, eo) and np.array_equal(bh, bo)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(spline_eval(host, masked, kh, eh, bh, sd, party, left, right, d, n, m, es),
                              spline_eval(OR, masked, ko, eo, bo, sd, party, left, right, d, n, m, es))


def test_ottt(host):
    rng = np.random.default_rng(8)
    N, n, es = 257, 10, 4
    alpha, x = rand_elems(rng, N, n, es), rand_elems(rng, N, n, es)
    masked = _masked(x, alpha, n, es)
    lut = rand_elems(rng, 2 << n, 32, es)
    key0 = rng.integers(0, 256, size=N * (1 << n) // 8, dtype=np.uint8)
    kh = ottt_keygen(host, key0, alpha, n, es)
    ko = ottt_keygen(OR, key0, alpha, n, es)
    assert np.array_equal(kh, ko)
    for party, k in ((0, key0), (1, ko)):
        eh, th = ottt_eval(host, masked, k, party, lut, n, es)
        eo, to = ottt_eval(OR, masked, k, party, lut, n, es)
        assert np.array_equal(eh, eo) and np.array_equal(th, to)


def test_errors_and_empty(host):
    """Validation runs in the device entry point behind the staging: the same codes as the oracle's;
    a NULL host buffer with a non-zero size is refused before anything runs."""
    z = np.zeros(4096, dtype=np.uint8)
    ks = 16 * 8 + 1 * 8 + 3
    for args in [(z, 1, z, 1, z, ks + 1, z, 16, 0, 8, 8, 1, 1, 1, None, 0),   # key size
                 (z, 1, z, 1, z, ks, z, 16, 2, 8, 8, 1, 1, 1, None, 0),       # party
                 (z, 1, z, 1, z, ks, z, 16, 0, 9, 8, 1, 1, 1, None, 0),       # bitWidthIn
                 (z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 3, 1, None, 0)]:      # element size
        assert host.call("dcfEval", *args) == OR.call("dcfEval", *args), args
    assert host.call("dcfEval", None, 0, None, 0, None, 0, None, 0, 0, 8, 8, 1, 1, 0, None, 0) == 0  # N = 0
    assert host.call("dcfEval", None, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0) == -1


def test_prng(host):
    lib, olib = host.lib, oracle()
    seed, ctr = bytes(range(40, 56)), le_block(2**64 - 2, 3)
    g, o = lib.FastFss_cpu_prngInit(), olib.oracle_prngInit()
    assert g and o
    try:
        s, c = ctypes.create_string_buffer(16), ctypes.create_string_buffer(16)
        assert lib.FastFss_cpu_prngGetCurrentSeed(g, s, c) == 0 and s.raw == bytes(16) and c.raw == bytes(16)
        assert lib.FastFss_cpu_prngSetCurrentSeed(g, seed, ctr) == 0
        assert olib.oracle_prngSetCurrentSeed(o, seed, ctr) == 0
        for es, num in ((4, 1001), (8, 3), (1, 17)):
            got, want = np.zeros(es * num, dtype=np.uint8), np.zeros(es * num, dtype=np.uint8)
            assert lib.FastFss_cpu_prngGen(g, got.ctypes.data, 8 * es, es, num) == 0
            assert olib.oracle_prngGen(o, want.ctypes.data, 8 * es, es, num) == 0
            assert np.array_equal(got, want), (es, num)
        c2 = ctypes.create_string_buffer(16)
        assert lib.FastFss_cpu_prngGetCurrentSeed(g, s, c) == 0
        olib.oracle_prngGetCurrentSeed(o, s, c2)
        assert c.raw == c2.raw and s.raw == seed
        assert lib.FastFss_cpu_prngGen(g, None, 8, 1, 4) == -17
        assert lib.FastFss_cpu_prngGen(g, got.ctypes.data, 9, 1, 4) == -18
    finally:
        lib.FastFss_cpu_prngRelease(g)
        olib.oracle_prngRelease(o)


def test_pyfastfss_cpu_tensors_equal_gpu_tensors():
    import torch

    import pyFastFss

    N, n = 1000, 20
    gen = torch.Generator().manual_seed(3)
    alpha = torch.randint(0, 1 << n, (N,), dtype=torch.int32, generator=gen)
    x = torch.randint(0, 1 << n, (N,), dtype=torch.int32, generator=gen)
    s0 = torch.randint(0, 256, (16 * N,), dtype=torch.uint8, generator=gen)
    s1 = torch.randint(0, 256, (16 * N,), dtype=torch.uint8, generator=gen)
    beta = torch.ones(N, dtype=torch.int32)
    outs = {}
    for dev in ("cpu", "cuda"):
        t = [v.to(dev) for v in (alpha, x, s0, s1, beta)]
        key = torch.empty(0, dtype=torch.uint8, device=dev)
        pyFastFss.dcf_key_gen(key, t[0], t[4], t[2], t[3], n, n, 1)
        y = torch.empty(0, dtype=torch.int32, device=dev)
        pyFastFss.dcf_eval(y, t[1], key, t[2], 0, n, n, 1)
        r = torch.empty(N, dtype=torch.int32, device=dev)
        p = pyFastFss.Prng(dev)
        p.set_current_seed(bytes(range(16)), bytes(16))
        p.rand_(r, 32)
        outs[dev] = (key.cpu(), y.cpu(), r.cpu())
    for a, b in zip(outs["cpu"], outs["cuda"]):
        assert torch.equal(a, b)
