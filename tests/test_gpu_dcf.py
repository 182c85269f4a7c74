"""DCF parity on the GPU: the device library against the oracle, bit-exact.

Small and medium sizes compare key bytes and both parties' shares with the
oracle element for element.  The 2^24-element n=32 configuration compares a
random sample of elements with the oracle and checks the size-independent
reconstruction property share0 + share1 == [x < alpha] on every element.
"""
import itertools

import numpy as np
import pytest

from fss_testlib import (DT, DeviceBackend, OracleBackend, dcf_eval, dcf_keygen, mask, ocall, rand_elems,
                         rand_seeds)

pytestmark = pytest.mark.gpu

OR = OracleBackend()


@pytest.fixture(scope="module")
def dev():
    return DeviceBackend()


GRID = {1: [1, 7, 8], 2: [1, 15, 16], 4: [1, 31, 32], 8: [1, 63, 64]}


def _case(rng, N, n, m, g, es):
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    beta = rand_elems(rng, N * g, m, es)
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    return alpha, beta, masked, rand_seeds(rng, N), rand_seeds(rng, N)


def _check(dev, alpha, beta, masked, s0, s1, n, m, g, es):
    key_o = dcf_keygen(OR, alpha, beta, s0, s1, n, m, g, es)
    key_d = dcf_keygen(dev, alpha, beta, s0, s1, n, m, g, es)
    assert np.array_equal(dev.get(key_d), key_o), "keyGen bytes differ"
    for party, seed in ((0, s0), (1, s1)):
        y_o = dcf_eval(OR, masked, key_o, seed, party, n, m, g, es)
        y_d = dev.get(dcf_eval(dev, masked, key_d, seed, party, n, m, g, es))
        assert np.array_equal(y_d, y_o), f"party {party} shares differ"


@pytest.mark.parametrize("es", [1, 2, 4, 8])
def test_dcf_reference_grid(dev, es):
    rng = np.random.default_rng(1000 + es)
    for n, m in itertools.product(GRID[es], GRID[es]):
        for g, N in [(1, 1023), (1, 1), (2, 8), (3, 33), (4, 2047)]:
            _check(dev, *_case(rng, N, n, m, g, es), n, m, g, es)


@pytest.mark.parametrize("n,m,g,es", [(24, 24, 6, 4), (32, 32, 5, 4), (16, 16, 9, 2), (24, 32, 48, 4),
                                      (64, 64, 3, 8)])
def test_dcf_aes_conversion_groups(dev, n, m, g, es):
    """groupSize*bytes > 16 takes the reference's AES refill conversion."""
    rng = np.random.default_rng(n * g)
    _check(dev, *_case(rng, 300, n, m, g, es), n, m, g, es)


def test_dcf_beta_null(dev):
    rng = np.random.default_rng(9)
    alpha, _, masked, s0, s1 = _case(rng, 500, 16, 16, 1, 2)
    _check(dev, alpha, None, masked, s0, s1, 16, 16, 1, 2)


def test_dcf_empty_and_errors(dev):
    import torch

    z = torch.zeros(4096, dtype=torch.uint8, device="cuda:0")
    zn = np.zeros(4096, dtype=np.uint8)
    ks = 16 * 8 + 8 + 3
    cases = [
        (z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 3, 1, None, 0),
        (z, 1, z, 1, z, ks, z, 16, 0, 9, 8, 1, 1, 1, None, 0),
        (z, 1, z, 1, z, ks, z, 16, 2, 8, 8, 1, 1, 1, None, 0),
        (z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 0, 1, 1, None, 0),
        (z, 1, z, 1, z, ks + 1, z, 16, 0, 8, 8, 1, 1, 1, None, 0),
        (z, 2, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0),
        (z, 1, z, 2, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0),
        (z, 1, z, 1, z, ks, z, 15, 0, 8, 8, 1, 1, 1, None, 0),
        (z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, z, 7),
    ]
    for c in cases:
        o = tuple(zn if (hasattr(a, "data_ptr")) else a for a in c)
        assert dev.call("dcfEval", *c) == ocall("dcfEval", *o) != 0
    # zero elements: success, nothing launched
    assert dev.call("dcfEval", None, 0, None, 0, None, 0, None, 0, 0, 8, 8, 1, 1, 0, None, 0) == 0


def _sub_key(key, idx, n, g, es, N):
    """Key bytes of the elements `idx` as a standalone len(idx)-element key."""
    sc = key[:N * 16 * n].reshape(N, 16 * n)[idx]
    tl = key[N * 16 * n:].reshape(N, es * (n * g + 2 + g))[idx]
    return np.concatenate([sc.reshape(-1), tl.reshape(-1)])


@pytest.mark.parametrize("es,n", [(1, 8), (1, 3), (2, 13), (4, 32), (4, 27), (4, 5), (8, 64)])
@pytest.mark.parametrize("N", [5000, 70001])
def test_eval_variants_agree(dev, es, n, N):
    """Every groupSize-1 launch variant (static bursts, elements per lane, the
    work queue, preloaded value corrections) gives the oracle's shares; 70001
    elements end in a partial tile, and n = 3, 5, 13, 27 end in a partial
    8-level burst."""
    rng = np.random.default_rng(es * 1000 + N)
    alpha, beta, masked, s0, s1 = _case(rng, N, n, n, 1, es)
    key_o = dcf_keygen(OR, alpha, beta, s0, s1, n, n, 1, es)
    key_d = dev.put(key_o)
    want = dcf_eval(OR, masked, key_o, s1, 1, n, n, 1, es)
    lib = dev.lib
    try:
        for v in range(lib.FastFss_hip_getEvalVariantCount()):
            assert lib.FastFss_hip_setEvalVariant(v) == 0
            got = dev.get(dcf_eval(dev, masked, key_d, s1, 1, n, n, 1, es))
            assert np.array_equal(got, want), v
    finally:
        lib.FastFss_hip_setEvalVariant(-1)
    assert lib.FastFss_hip_setEvalVariant(lib.FastFss_hip_getEvalVariantCount()) == -1


def test_dcf_cpu_workload_n16_2pow16(dev):
    """The reference's CPU-sized workload (n = m = 16 in 64-bit elements, 2^16
    elements, beta = NULL) on the GPU, bit-exact against the oracle."""
    rng = np.random.default_rng(16)
    N, n, m, es = 1 << 16, 16, 16, 8
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    _check(dev, alpha, None, x, s0, s1, n, m, 1, es)


def full_size_chunks(N, sk, tail, key_h, gen_key, evals, chunk=1 << 22, what=""):
    """Compares a full-size device key set and its shares with the oracle, chunk by chunk.

    key_h is the device key (bytes, layout [sCW: N x sk][tail: N x tail]); gen_key(sl) makes the
    oracle's key for the elements of slice sl, and evals is a list of (oracle_eval(sl, key), device
    shares) pairs."""
    for c0 in range(0, N, chunk):
        c1 = min(N, c0 + chunk)
        sl = slice(c0, c1)
        ko = gen_key(sl)
        kd = np.concatenate([key_h[sk * c0:sk * c1], key_h[N * sk + tail * c0:N * sk + tail * c1]])
        assert np.array_equal(ko, kd), f"{what} key bytes differ in elements [{c0}, {c1})"
        for j, (ev, got) in enumerate(evals):
            assert np.array_equal(ev(sl, ko), got[sl]), f"{what} party {j} shares differ in [{c0}, {c1})"
        print(f"{what} elements [{c0}, {c1}): keys and both parties' shares equal the oracle's", flush=True)


def test_dcf_2pow24_n32(dev):
    """Plain DCF, bitWidthIn = bitWidthOut = 32, 2^24 elements, one key each: every key byte and both
    parties' shares over all 2^24 elements equal the oracle's, and share0 + share1 == [x < alpha]."""
    import torch

    N, n, m, g, es = 1 << 24, 32, 32, 1, 4
    gen = torch.Generator(device="cuda:0").manual_seed(2024)
    alpha = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    x = torch.randint(0, 2**32, (N,), dtype=torch.int64, device="cuda:0", generator=gen)
    masked = (x + alpha) & 0xFFFFFFFF
    seeds = torch.randint(0, 256, (2, 16 * N), dtype=torch.uint8, device="cuda:0", generator=gen)
    a32 = alpha.to(torch.int32)  # bit patterns of the uint32 elements
    m32 = masked.to(torch.int32)
    key = dcf_keygen(dev, a32, None, seeds[0], seeds[1], n, m, g, es)
    y0 = dcf_eval(dev, m32, key, seeds[0], 0, n, m, g, es)
    y1 = dcf_eval(dev, m32, key, seeds[1], 1, n, m, g, es)
    rec = (y0.to(torch.int64) + y1.to(torch.int64)) & 0xFFFFFFFF
    want = (masked < alpha).to(torch.int64)
    assert torch.equal(rec, want)
    del rec, want, x, alpha, masked
    key_h, a_h, m_h = dev.get(key), dev.get(a32), dev.get(m32)
    s0_h, s1_h = dev.get(seeds[0]), dev.get(seeds[1])
    y0_h, y1_h = dev.get(y0), dev.get(y1)
    del key
    sd = lambda s, sl: s[16 * sl.start:16 * sl.stop]
    full_size_chunks(N, 16 * n, (n * g + 2 + g) * es, key_h,
                     lambda sl: dcf_keygen(OR, a_h[sl], None, sd(s0_h, sl), sd(s1_h, sl), n, m, g, es),
                     [(lambda sl, ko: dcf_eval(OR, m_h[sl], ko, sd(s0_h, sl), 0, n, m, g, es), y0_h),
                      (lambda sl, ko: dcf_eval(OR, m_h[sl], ko, sd(s1_h, sl), 1, n, m, g, es), y1_h)], what="DCF")
