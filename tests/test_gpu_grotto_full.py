"""Full-domain Grotto (FastFss_hip_grottoEvalFullDomain) on the GPU.

The entry point is defined point for point by the reference's per-point Grotto
evaluations, so every case compares it with the oracle's grottoEval /
grottoEqEval run on every x of the domain (keys replicated per point), for
both output layouts and every subtree split down to depth-first expansion
of 8 levels per lane.  At the full workload size (4096 keys, n = 16) the two
parties' packed shares are checked to XOR to [x < alpha] / [x <= alpha] /
[x == alpha] for every key and point, and a sample of keys against the
device's per-point evaluation.
"""
import ctypes

import numpy as np
import pytest

from fss_testlib import (DT, DeviceBackend, OracleBackend, grotto_eq_eval, grotto_eval, grotto_keygen, nbytes,
                         ptr, rand_elems, rand_seeds)

pytestmark = pytest.mark.gpu

OR = OracleBackend()
LT, LE, EQ = 0, 1, 2


@pytest.fixture(scope="module")
def dev():
    d = DeviceBackend()
    yield d
    d.lib.FastFss_hip_grottoFullSetLaneTarget(0)


def _replicate(key, n, es, N, idx):
    """Grotto key whose element j is element idx[j] of `key` (layout src/impl/grotto.h:53-79)."""
    lv = n - 6
    sc = key[:N * 16 * lv].reshape(N, 16 * lv)
    last = key[N * 16 * lv:N * 16 * lv + 8 * N].reshape(N, 8)
    tlr = key[N * 16 * lv + 8 * N:].reshape(N, 2 * es)
    return np.concatenate([sc[idx].reshape(-1), last[idx].reshape(-1), tlr[idx].reshape(-1)])


def _oracle_full(key, seeds, party, n, es, N, mode):
    idx = np.repeat(np.arange(N), 1 << n)
    k = _replicate(key, n, es, N, idx)
    s = seeds.reshape(N, 16)[idx].reshape(-1)
    x = np.tile(np.arange(1 << n, dtype=np.uint64), N).astype(DT[es])
    if mode == EQ:
        return grotto_eq_eval(OR, x, k, s, party, n, es)
    return grotto_eval(OR, x, k, s, party, n, es, mode == LE)


def _full(dev, key, seeds, party, n, es, N, mode, packed):
    torch = dev.torch
    words = N << (n - 6)
    out = dev.empty(8 * words if packed else (N << n) * es, np.uint64 if packed else DT[es])
    sd = dev.put(seeds)
    rc = dev.lib.FastFss_hip_grottoEvalFullDomain(ptr(out), nbytes(out), ptr(key), nbytes(key), ptr(sd), nbytes(sd),
                                                  party, mode, packed, n, es, N, ctypes.addressof(dev._stream))
    assert rc == 0, rc
    return out


def _unpack(words):
    return np.unpackbits(words.view(np.uint8), bitorder="little")


@pytest.mark.parametrize("n,es,N,lanes", [(6, 1, 5, 0), (8, 1, 3, 1), (10, 2, 4, 16), (12, 2, 3, 1), (14, 4, 2, 1),
                                          (13, 8, 3, 4), (12, 4, 7, 0)])
def test_full_domain_matches_per_point_oracle(dev, n, es, N, lanes):
    rng = np.random.default_rng(n * 31 + es)
    alpha = rand_elems(rng, N, n, es)
    alpha[0] = 0
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ko = grotto_keygen(OR, alpha, s0, s1, n, es)
    kd = dev.put(ko)
    dev.lib.FastFss_hip_grottoFullSetLaneTarget(lanes)  # small targets force deep depth-first subtrees
    try:
        for party, sd in ((0, s0), (1, s1)):
            for mode in (LT, LE, EQ):
                want = _oracle_full(ko, sd, party, n, es, N, mode)
                got = dev.get(_full(dev, kd, sd, party, n, es, N, mode, 0))
                assert np.array_equal(got, want), (party, mode)
                packed = dev.get(_full(dev, kd, sd, party, n, es, N, mode, 1))
                assert np.array_equal(_unpack(packed), want.astype(np.uint8)), (party, mode, "packed")
    finally:
        dev.lib.FastFss_hip_grottoFullSetLaneTarget(0)


def test_full_domain_errors(dev):
    n, es, N = 10, 2, 4
    rng = np.random.default_rng(1)
    s0 = rand_seeds(rng, N)
    key = dev.put(grotto_keygen(OR, rand_elems(rng, N, n, es), s0, s0, n, es))
    sd = dev.put(s0)
    out = dev.empty((N << n) * es)
    call = dev.lib.FastFss_hip_grottoEvalFullDomain
    sp = ctypes.addressof(dev._stream)
    ok = (ptr(out), nbytes(out), ptr(key), nbytes(key), ptr(sd), nbytes(sd))
    assert call(*ok, 0, LT, 0, n, es, N, sp) == 0
    assert call(*ok, 2, LT, 0, n, es, N, sp) == -12
    assert call(*ok, 0, 3, 0, n, es, N, sp) == -1
    assert call(*ok, 0, LT, 2, n, es, N, sp) == -1
    assert call(*ok, 0, LT, 0, 5, es, N, sp) == -10
    assert call(*ok, 0, LT, 0, 17, es, N, sp) == -10
    assert call(*ok, 0, LT, 0, n, 3, N, sp) == -11
    assert call(ok[0], ok[1] - 2, *ok[2:], 0, LT, 0, n, es, N, sp) == -8
    assert call(ok[0], ok[1], ok[2], ok[3] - 1, *ok[4:], 0, LT, 0, n, es, N, sp) == -2
    assert call(*ok[:4], ok[4], ok[5] - 16, 0, LT, 0, n, es, N, sp) == -5
    assert call(*ok, 0, LT, 1, n, es, N, sp) == -8  # packed layout is smaller
    assert call(None, 0, None, 0, None, 0, 0, LT, 0, n, es, 0, sp) == 0


@pytest.mark.parametrize("N", [4096, 4095])
def test_full_domain_4096_keys_n16(dev, N):
    """Grotto at bitWidthIn = 16 over 4096 keys, full domain 2^16 points.  At 4095 keys each lane
    owns a subtree of depth 3 below 7 shared top levels, and a workgroup makes two passes whose
    key groups straddle workgroup boundaries."""
    torch = dev.torch
    n, es = 16, 2
    rng = np.random.default_rng(N)
    alpha = rand_elems(rng, N, n, es)
    alpha[:3] = [0, 1, 0xFFFF]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kd = grotto_keygen(dev, alpha, s0, s1, n, es)
    a = torch.from_numpy(alpha.astype(np.int64)).to("cuda:0")
    words = torch.arange(1 << (n - 6), device="cuda:0", dtype=torch.int64)
    lo = (words * 64)[None, :]  # first point of each word
    p = (a[:, None] - lo)  # position of alpha inside the word, if there
    full = torch.full_like(p, -1)
    bit = torch.ones_like(p)
    inside = (p >= 0) & (p < 64)
    sh = p.clamp(0, 63)
    lt_want = torch.where(p >= 64, full, torch.where(inside, (bit << sh) - 1, torch.zeros_like(p)))
    eq_want = torch.where(inside, bit << sh, torch.zeros_like(p))
    want = {LT: lt_want, EQ: eq_want, LE: lt_want ^ eq_want}
    for mode in (LT, LE, EQ):
        y0 = _full(dev, kd, s0, 0, n, es, N, mode, 1).view(torch.int64).reshape(N, -1)
        y1 = _full(dev, kd, s1, 1, n, es, N, mode, 1).view(torch.int64).reshape(N, -1)
        assert torch.equal(y0 ^ y1, want[mode]), mode
    # element layout on a sample of keys equals the per-point device evaluation
    ge = _full(dev, kd, s0, 0, n, es, N, LT, 0).reshape(N, 1 << n)
    key_h = dev.get(kd)
    for e in (0, 1, 2, 1234, N - 1):
        idx = np.full(1 << n, e)
        k1 = dev.put(_replicate(key_h, n, es, N, idx))
        s = s0.reshape(N, 16)[idx].reshape(-1)
        x = np.arange(1 << n, dtype=np.uint16)
        per_point = grotto_eval(dev, x, k1, s, 0, n, es, False)
        assert torch.equal(ge[e], per_point), e
