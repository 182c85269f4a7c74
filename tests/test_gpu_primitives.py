"""DPF, Grotto, MIC, spline, OTTT and PRNG parity on the GPU against the oracle.

Every case compares the device library's key bytes and both parties' shares
with the oracle bit for bit on the same inputs, and checks the reconstruction
property the reference's own tests assert.
"""
import ctypes
import itertools

import numpy as np
import pytest

from fss_testlib import (DT, DeviceBackend, OracleBackend, dpf_eval, dpf_eval_all, dpf_eval_multi, dpf_keygen,
                         grotto_eq_eval, grotto_eval, grotto_keygen, grotto_lut_eval, grotto_mic_eval, le_block, mask,
                         mic_eval, mic_keygen, oracle, ottt_eval, ottt_keygen, rand_elems, rand_seeds, signed_nbit,
                         spline_eval, spline_keygen)

pytestmark = pytest.mark.gpu

OR = OracleBackend()


@pytest.fixture(scope="module")
def dev():
    return DeviceBackend()


def _masked(x, alpha, n, es):
    return ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])


GRID = {1: [1, 7, 8], 2: [1, 15, 16], 4: [1, 31, 32], 8: [1, 63, 64]}


# ------------------------------------------------------------------ DPF ----


@pytest.mark.parametrize("es", [1, 2, 4, 8])
def test_dpf_grid(dev, es):
    rng = np.random.default_rng(200 + es)
    for n, m in itertools.product(GRID[es], GRID[es]):
        for g, N in [(1, 777), (2, 16), (3, 1)]:
            alpha = rand_elems(rng, N, n, es)
            x = rand_elems(rng, N, n, es)
            x[::3] = 0
            beta = rand_elems(rng, N * g, m, es)
            masked = _masked(x, alpha, n, es)
            s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
            ko = dpf_keygen(OR, alpha, beta, s0, s1, n, m, g, es)
            kd = dpf_keygen(dev, alpha, beta, s0, s1, n, m, g, es)
            assert np.array_equal(dev.get(kd), ko), (n, m, g)
            for party, sd in ((0, s0), (1, s1)):
                yo = dpf_eval(OR, masked, ko, sd, party, n, m, g, es)
                yd = dev.get(dpf_eval(dev, masked, kd, sd, party, n, m, g, es))
                assert np.array_equal(yd, yo), (n, m, g, party)


@pytest.mark.parametrize("n,es,g", [(1, 1, 1), (4, 1, 2), (10, 2, 1), (8, 4, 3), (12, 8, 1)])
def test_dpf_eval_all_and_multi(dev, n, es, g):
    rng = np.random.default_rng(n * 31 + g)
    N, m = 37, 8 * es
    alpha = rand_elems(rng, N, n, es)
    beta = rand_elems(rng, N * g, m, es)
    x = rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko = dpf_keygen(OR, alpha, beta, s0, s1, n, m, g, es)
    kd = dpf_keygen(dev, alpha, beta, s0, s1, n, m, g, es)
    pts = rand_elems(rng, 9, 8 * es, es)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dev.get(dpf_eval_all(dev, x, kd, sd, party, n, m, g, es)),
                              dpf_eval_all(OR, x, ko, sd, party, n, m, g, es))
        assert np.array_equal(dev.get(dpf_eval_multi(dev, x, kd, sd, party, pts, n, m, g, es)),
                              dpf_eval_multi(OR, x, ko, sd, party, pts, n, m, g, es))


@pytest.mark.parametrize("n,m,es,g,lanes", [(10, 8, 2, 1, 1), (10, 32, 4, 1, 16), (12, 16, 2, 1, 4),
                                             (9, 20, 4, 3, 2), (14, 64, 8, 1, 1)])
def test_dpf_eval_all_subtrees(dev, n, m, es, g, lanes):
    """evalAll expands subtrees depth-first; small lane targets force deep
    subtrees (up to 8 levels) so every stack depth is exercised."""
    rng = np.random.default_rng(n * 7 + m + g)
    N = 3
    alpha = rand_elems(rng, N, n, es)
    beta = rand_elems(rng, N * g, m, es)
    x = rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko = dpf_keygen(OR, alpha, beta, s0, s1, n, m, g, es)
    kd = dev.put(ko)
    dev.lib.FastFss_hip_grottoFullSetLaneTarget(lanes)
    try:
        for party, sd in ((0, s0), (1, s1)):
            assert np.array_equal(dev.get(dpf_eval_all(dev, x, kd, sd, party, n, m, g, es)),
                                  dpf_eval_all(OR, x, ko, sd, party, n, m, g, es)), party
    finally:
        dev.lib.FastFss_hip_grottoFullSetLaneTarget(0)


@pytest.mark.parametrize("es,n,m", [(2, 16, 8), (4, 32, 32), (8, 64, 20)])
def test_dpf_queue_size(dev, es, n, m):
    """70001 elements: the static split with a partial last stride (the work
    queue starts at 2^21 elements, test_dpf_queue_2pow21)."""
    rng = np.random.default_rng(es + n)
    N = 70001
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    x[::5] = alpha[::5]
    beta = rand_elems(rng, N, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko = dpf_keygen(OR, alpha, beta, s0, s1, n, m, 1, es)
    kd = dpf_keygen(dev, alpha, beta, s0, s1, n, m, 1, es)
    assert np.array_equal(dev.get(kd), ko)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dev.get(dpf_eval(dev, x, kd, sd, party, n, m, 1, es)),
                              dpf_eval(OR, x, ko, sd, party, n, m, 1, es)), party


def test_dpf_2pow24_n32(dev):
    """Plain DPF, bitWidthIn = bitWidthOut = 32, 2^24 elements, one key each, random beta: every key
    byte and both parties' shares over all 2^24 elements equal the oracle's, and share0 + share1 ==
    beta * [x == alpha] (x == alpha on every fifth element)."""
    import torch

    from test_gpu_dcf import full_size_chunks

    N, n, m, es = 1 << 24, 32, 32, 4
    gen = torch.Generator(device="cuda:0").manual_seed(24)
    alpha = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    x = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    x[::5] = alpha[::5]
    beta = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    seeds = torch.randint(0, 256, (2, 16 * N), dtype=torch.uint8, device="cuda:0", generator=gen)
    a32, x32, b32 = alpha.to(torch.int32), x.to(torch.int32), beta.to(torch.int32)
    key = dpf_keygen(dev, a32, b32, seeds[0], seeds[1], n, m, 1, es)
    y0 = dpf_eval(dev, x32, key, seeds[0], 0, n, m, 1, es)
    y1 = dpf_eval(dev, x32, key, seeds[1], 1, n, m, 1, es)
    rec = (y0.to(torch.int64) + y1.to(torch.int64)) & 0xFFFFFFFF
    assert torch.equal(rec, torch.where(x == alpha, beta, torch.zeros_like(beta)))
    del rec, alpha, x, beta
    key_h, a_h, x_h, b_h = dev.get(key), dev.get(a32), dev.get(x32), dev.get(b32)
    s0_h, s1_h = dev.get(seeds[0]), dev.get(seeds[1])
    y0_h, y1_h = dev.get(y0), dev.get(y1)
    del key
    sd = lambda s, sl: s[16 * sl.start:16 * sl.stop]
    full_size_chunks(N, 16 * n, 3 * es, key_h,
                     lambda sl: dpf_keygen(OR, a_h[sl], b_h[sl], sd(s0_h, sl), sd(s1_h, sl), n, m, 1, es),
                     [(lambda sl, ko: dpf_eval(OR, x_h[sl], ko, sd(s0_h, sl), 0, n, m, 1, es), y0_h),
                      (lambda sl, ko: dpf_eval(OR, x_h[sl], ko, sd(s1_h, sl), 1, n, m, 1, es), y1_h)], what="DPF")


def test_dpf_queue_2pow21(dev):
    """From 2^21 elements the groupSize-1 DPF eval runs on the work queue: 2^21 + 77
    elements (n = m = 32) end in a partial tile.  Every element reconstructs to
    beta * [x == alpha]; a sample of 2048 keys and shares equals the oracle's."""
    import torch

    N, n, m, es = (1 << 21) + 77, 32, 32, 4
    gen = torch.Generator(device="cuda:0").manual_seed(21)
    alpha = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    x = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    x[::5] = alpha[::5]
    beta = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    seeds = torch.randint(0, 256, (2, 16 * N), dtype=torch.uint8, device="cuda:0", generator=gen)
    a32, x32, b32 = alpha.to(torch.int32), x.to(torch.int32), beta.to(torch.int32)
    key = dpf_keygen(dev, a32, b32, seeds[0], seeds[1], n, m, 1, es)
    y0 = dpf_eval(dev, x32, key, seeds[0], 0, n, m, 1, es)
    y1 = dpf_eval(dev, x32, key, seeds[1], 1, n, m, 1, es)
    rec = (y0.to(torch.int64) + y1.to(torch.int64)) & 0xFFFFFFFF
    assert torch.equal(rec, torch.where(x == alpha, beta, torch.zeros_like(beta)))
    idx = np.sort(np.random.default_rng(7).choice(N, 2048, replace=False))
    kh = dev.get(key)
    sub = np.concatenate([kh[:N * 16 * n].reshape(N, 16 * n)[idx].reshape(-1),
                          kh[N * 16 * n:].reshape(N, es * 3)[idx].reshape(-1)])
    s0 = dev.get(seeds[0]).reshape(N, 16)[idx].reshape(-1)
    s1 = dev.get(seeds[1]).reshape(N, 16)[idx].reshape(-1)
    ko = dpf_keygen(OR, dev.get(a32)[idx], dev.get(b32)[idx], s0, s1, n, m, 1, es)
    assert np.array_equal(ko, sub)
    xh = dev.get(x32)[idx]
    assert np.array_equal(dpf_eval(OR, xh, ko, s0, 0, n, m, 1, es), dev.get(y0)[idx])
    assert np.array_equal(dpf_eval(OR, xh, ko, s1, 1, n, m, 1, es), dev.get(y1)[idx])


# --------------------------------------------------------------- Grotto ----


@pytest.mark.parametrize("n,es", [(6, 1), (8, 1), (16, 2), (24, 4), (32, 4), (64, 8), (40, 8)])
def test_grotto(dev, n, es):
    rng = np.random.default_rng(n + es)
    N = 3000
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    x[::4] = 0
    x[1::7] = rand_elems(rng, len(x[1::7]), min(n, 6), es)  # small offsets exercise the equalBound early exit
    masked = _masked(x, alpha, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko = grotto_keygen(OR, alpha, s0, s1, n, es)
    kd = grotto_keygen(dev, alpha, s0, s1, n, es)
    assert np.array_equal(dev.get(kd), ko)
    for party, sd in ((0, s0), (1, s1)):
        for eb in (False, True):
            assert np.array_equal(dev.get(grotto_eval(dev, masked, kd, sd, party, n, es, eb)),
                                  grotto_eval(OR, masked, ko, sd, party, n, es, eb)), (party, eb)
        assert np.array_equal(dev.get(grotto_eq_eval(dev, masked, kd, sd, party, n, es)),
                              grotto_eq_eval(OR, masked, ko, sd, party, n, es))


def _intervals(rng, I, n, es):
    lo = rand_elems(rng, I, n, es)
    hi = rand_elems(rng, I, n, es)
    L, R = [], []
    for a, b in zip(lo, hi):
        a, b = int(a), int(b)
        if signed_nbit(b, n) < signed_nbit(a, n):
            a, b = b, a
        L.append(a)
        R.append(b)
    L[min(2, I - 1)] = (R[min(1, I - 1)] + 1) & mask(n)
    return np.array(L, dtype=DT[es]), np.array(R, dtype=DT[es])


@pytest.mark.parametrize("n,es,I", [(8, 1, 3), (16, 2, 5), (24, 4, 16), (64, 8, 4)])
def test_grotto_mic_lut(dev, n, es, I):
    rng = np.random.default_rng(n * I)
    N = 500
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    masked = _masked(x, alpha, n, es)
    left, right = _intervals(rng, I, n, es)
    lut = rand_elems(rng, 3 * I, 8 * es, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko = grotto_keygen(OR, alpha, s0, s1, n, es)
    kd = grotto_keygen(dev, alpha, s0, s1, n, es)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dev.get(grotto_mic_eval(dev, masked, kd, sd, party, left, right, n, es)),
                              grotto_mic_eval(OR, masked, ko, sd, party, left, right, n, es))
        de, dt = grotto_lut_eval(dev, masked, kd, sd, party, left, right, lut, n, 8 * es, es)
        oe, ot = grotto_lut_eval(OR, masked, ko, sd, party, left, right, lut, n, 8 * es, es)
        assert np.array_equal(dev.get(de), oe) and np.array_equal(dev.get(dt), ot)


# ------------------------------------------------------------------ MIC ----


@pytest.mark.parametrize("n,m,es,I,N", [(8, 8, 1, 3, 400), (16, 16, 2, 5, 300), (24, 32, 4, 16, 2000),
                                        (31, 7, 4, 6, 100), (64, 64, 8, 4, 200), (24, 32, 4, 16, 3)])
def test_mic(dev, n, m, es, I, N):
    rng = np.random.default_rng(n * m + I)
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    masked = _masked(x, alpha, n, es)
    left, right = _intervals(rng, I, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko, zo = mic_keygen(OR, alpha, s0, s1, left, right, n, m, es)
    kd, zd = mic_keygen(dev, alpha, s0, s1, left, right, n, m, es)
    assert np.array_equal(dev.get(kd), ko) and np.array_equal(dev.get(zd), zo)
    z0 = rand_elems(rng, N * I, 8 * es, es)
    z1 = ((zo.astype(np.uint64) - z0) & np.uint64(mask(8 * es))).astype(DT[es])
    outs = []
    for party, sd, zz in ((0, s0, z0), (1, s1, z1)):
        yo = mic_eval(OR, masked, ko, zz, sd, party, left, right, n, m, es)
        yd = dev.get(mic_eval(dev, masked, kd, zz, sd, party, left, right, n, m, es))
        assert np.array_equal(yd, yo), party
        outs.append(yd.astype(np.uint64))
    rec = ((outs[0] + outs[1]) & np.uint64(mask(m))).reshape(N, I)
    for i in range(min(N, 50)):
        xs = signed_nbit(int(x[i]), n)
        for q in range(I):
            lo, hi = signed_nbit(int(left[q]), n), signed_nbit(int(right[q]), n)
            if lo <= hi:  # the forced adjacent interval may wrap; its parity is covered above
                assert int(rec[i, q]) == int(lo <= xs <= hi)


# --------------------------------------------------------------- spline ----


def _spline_case(rng, I, d, n, m, es):
    M = mask(n)
    r = sorted(int(v) for v in rand_elems(rng, I, n, es))
    for i in range(I - 1):
        lo = 0 if i == 0 else r[i - 1] + 1
        hi = M - (I - 1 - i)
        r[i] = lo + r[i] % (hi - lo + 1)
    r[-1] = M
    left = [0] + [v + 1 for v in r[:-1]]
    return np.array(left, dtype=DT[es]), np.array(r, dtype=DT[es]), rand_elems(rng, I * (d + 1), m, es)


@pytest.mark.parametrize("n,m,es,I,d", [(8, 8, 1, 4, 2), (16, 16, 2, 5, 1), (24, 32, 4, 16, 2), (32, 32, 4, 3, 7),
                                        (20, 20, 4, 2, 9), (48, 48, 8, 3, 2),
                                        # 3-byte and 5-byte outputs: slots start at unaligned byte offsets
                                        (24, 24, 4, 4, 2), (40, 40, 8, 3, 1), (12, 10, 2, 5, 4)])
def test_spline(dev, n, m, es, I, d):
    rng = np.random.default_rng(n + I * d)
    N = 300
    left, right, coeff = _spline_case(rng, I, d, n, m, es)
    if I > 3:  # one gap: the point x-1-L_q of a non-adjacent interval is walked on its own
        left = left.copy()
        left[3] = (int(left[3]) + 1) & mask(n)
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    masked = _masked(x, alpha, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko, eo, bo = spline_keygen(OR, alpha, s0, s1, coeff, d, left, right, n, m, es)
    kd, ed, bd = spline_keygen(dev, alpha, s0, s1, coeff, d, left, right, n, m, es)
    assert np.array_equal(dev.get(kd), ko)
    assert np.array_equal(dev.get(ed), eo) and np.array_equal(dev.get(bd), bo)
    W = np.uint64(mask(8 * es))
    b0 = rand_elems(rng, len(bo), 8 * es, es)
    e0 = rand_elems(rng, len(eo), 8 * es, es)
    b1 = ((bo.astype(np.uint64) - b0) & W).astype(DT[es])
    e1 = ((eo.astype(np.uint64) - e0) & W).astype(DT[es])
    for party, sd, bb, ee in ((0, s0, b0, e0), (1, s1, b1, e1)):
        yo = spline_eval(OR, masked, ko, ee, bb, sd, party, left, right, d, n, m, es)
        yd = dev.get(spline_eval(dev, masked, kd, ee, bb, sd, party, left, right, d, n, m, es))
        assert np.array_equal(yd, yo), party


# ----------------------------------------------------------------- OTTT ----


@pytest.mark.parametrize("n,es,lutNum", [(3, 1, 1), (4, 1, 2), (5, 1, 1), (8, 2, 3), (12, 4, 2), (9, 8, 1),
                                         # n = 6: one 64-bit word; tables beyond four: several passes
                                         (6, 1, 2), (10, 4, 5), (7, 8, 9),
                                         # tables staged in LDS chunks (80 KB budget): 4 chunks each;
                                         # n = 14 in 2-byte elements fits whole (64 KB)
                                         (13, 8, 4), (16, 4, 1), (14, 2, 0)])
def test_ottt(dev, n, es, lutNum):
    rng = np.random.default_rng(n * 5 + lutNum)
    N = 513
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    masked = _masked(x, alpha, n, es)
    lut = rand_elems(rng, lutNum << n, 8 * es, es)
    key0 = rng.integers(0, 256, size=N * (1 << n) // 8, dtype=np.uint8)
    k1o = ottt_keygen(OR, key0, alpha, n, es)
    k1d = ottt_keygen(dev, key0, alpha, n, es)
    assert np.array_equal(dev.get(k1d), k1o)
    for party, key in ((0, key0), (1, k1o)):
        oe, ot = ottt_eval(OR, masked, key, party, lut, n, es)
        de, dt = ottt_eval(dev, masked, dev.put(key), party, lut, n, es)
        assert np.array_equal(dev.get(de), oe) and np.array_equal(dev.get(dt), ot)


# ----------------------------------------------------------------- PRNG ----


@pytest.mark.parametrize("es,num", [(1, 37), (4, 4097), (8, 3), (2, 1 << 20)])
def test_prng(dev, es, num):
    import torch

    lib = dev.lib
    olib = oracle()
    seed = bytes(range(16, 32))
    ctr = le_block(2**64 - 3, 9)
    g = lib.FastFss_cuda_prngInit()
    o = olib.oracle_prngInit()
    assert g and o
    assert lib.FastFss_cuda_prngSetCurrentSeed(g, seed, ctr) == 0
    assert olib.oracle_prngSetCurrentSeed(o, seed, ctr) == 0
    for _ in range(2):  # the second call continues from the advanced counter
        dst = torch.zeros(es * num, dtype=torch.uint8, device="cuda:0")
        rc = lib.FastFss_cuda_prngGen(g, dst.data_ptr(), 8 * es, es, num, ctypes.addressof(dev._stream))
        assert rc == 0
        want = np.zeros(es * num, dtype=np.uint8)
        assert olib.oracle_prngGen(o, want.ctypes.data, 8 * es, es, num) == 0
        assert np.array_equal(dev.get(dst), want)
    s = ctypes.create_string_buffer(16)
    c1 = ctypes.create_string_buffer(16)
    c2 = ctypes.create_string_buffer(16)
    torch.cuda.synchronize()
    assert lib.FastFss_cuda_prngGetCurrentSeed(g, s, c1) == 0
    olib.oracle_prngGetCurrentSeed(o, s, c2)
    assert c1.raw == c2.raw and s.raw == seed
    assert lib.FastFss_cuda_prngGen(g, None, 8, 1, 1, None) == -17
    assert lib.FastFss_cuda_prngGen(g, dst.data_ptr(), 9, 1, 1, None) == -18
    lib.FastFss_cuda_prngRelease(g)
    olib.oracle_prngRelease(o)
