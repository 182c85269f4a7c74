"""FastFss_hip_dcfEvalNoAlloc against FastFss_cuda_dcfEval, bit for bit.

The no-allocation entry point keeps the work queue's tile counter in static
device memory shared by every call on the GPU, so the calls below run in a row
on one stream and each one must start from a cleared counter: a larger call
first, then smaller ones that would skip tiles if the counter kept the previous
call's value.  A sample of each call is also compared with the oracle.
"""
import ctypes

import numpy as np
import pytest

from fss_testlib import (DT, DeviceBackend, OracleBackend, dcf_eval, dcf_keygen, key_size, mask, nbytes, ptr,
                         rand_elems, rand_seeds)

pytestmark = pytest.mark.gpu

OR = OracleBackend()
MAX_ELEMS = 1 << 36  # FAST_FSS_HIP_NO_ALLOC_MAX_ELEMENTS


@pytest.fixture(scope="module")
def dev():
    return DeviceBackend()


def _no_alloc(dev, x, key, seed, party, n, m, g, es, N):
    out = dev.empty(N * es * g, DT[es])
    rc = dev.lib.FastFss_hip_dcfEvalNoAlloc(ptr(out), nbytes(out), ptr(x), nbytes(x), ptr(key), nbytes(key),
                                            ptr(seed), nbytes(seed), party, n, m, g, es, N, None, 0,
                                            ctypes.addressof(dev._stream))
    assert rc == 0, rc
    return out


def _run(dev, rng, N, n, m, g, es, party, variant=-1, sample=256):
    print(f"no-alloc DCF: N={N} n={n} m={m} g={g} es={es} party={party} variant={variant}", flush=True)
    alpha = rand_elems(rng, N, n, es)
    xs = rand_elems(rng, N, n, es)
    masked = ((xs.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dcf_keygen(dev, alpha, None if g == 1 else rand_elems(rng, N * g, m, es), s0, s1, n, m, g, es)
    seed = s0 if party == 0 else s1
    x_d, seed_d = dev.put(masked), dev.put(seed)
    assert dev.lib.FastFss_hip_setEvalVariant(variant) == 0
    try:
        got = dev.get(_no_alloc(dev, x_d, key, seed_d, party, n, m, g, es, N))
        want = dev.get(dcf_eval(dev, x_d, key, seed_d, party, n, m, g, es))
    finally:
        dev.lib.FastFss_hip_setEvalVariant(-1)
    assert np.array_equal(got, want), "no-alloc shares differ from FastFss_cuda_dcfEval"
    # a sample against the oracle on the device key's bytes
    idx = np.unique(np.concatenate([rng.choice(N, min(sample, N), replace=False), [0, N - 1]]))
    key_h = dev.get(key)
    ks = n * 16
    tail = (n * g + 2 + g) * es
    sub_key = np.concatenate([key_h[:N * ks].reshape(N, ks)[idx].ravel(),
                              key_h[N * ks:].reshape(N, tail)[idx].ravel()])
    sub_seed = seed.reshape(N, 16)[idx].ravel()
    y_o = dcf_eval(OR, masked[idx], sub_key, sub_seed, party, n, m, g, es)
    assert np.array_equal(got.reshape(N, g)[idx].ravel(), y_o), "no-alloc shares differ from the oracle"


def test_no_alloc_matches_dcf_eval(dev):
    rng = np.random.default_rng(77)
    # the work queue with preloaded corrections (the default from 2^20 elements), then a smaller call
    _run(dev, rng, (1 << 21) + 77, 32, 32, 1, 4, 0)
    _run(dev, rng, (1 << 20) + 3, 32, 32, 1, 4, 1)
    # the queue forced on small calls: small-workgroup kernels, 64-bit elements (no preload)
    _run(dev, rng, 5000, 32, 32, 1, 4, 0, variant=6)
    _run(dev, rng, 3001, 64, 64, 1, 8, 1, variant=6)
    _run(dev, rng, 700, 16, 16, 1, 2, 0, variant=5)
    # sizes and variants without the queue take the same kernels as FastFss_cuda_dcfEval
    _run(dev, rng, 1000, 24, 24, 1, 4, 1)
    _run(dev, rng, 333, 32, 32, 3, 4, 0)


def test_no_alloc_validation(dev):
    """Same codes as FastFss_cuda_dcfEval on invalid sizes; beyond the maximum element count the call
    returns FAST_FSS_RUNTIME_ERROR before touching any buffer."""
    lib = dev.lib
    st = ctypes.addressof(dev._stream)
    n, m, g, es = 32, 32, 1, 4
    # every case below is rejected by validation before any launch (no buffer is read)
    N = 10
    ks = key_size(dev, "dcfGetKeyDataSize", n, m, g, es, N)
    for args, code in [((None, N * es * g, None, N * es, None, ks + 1, None, N * 16, 0), -2),
                       ((None, N * es * g + 1, None, N * es, None, ks, None, N * 16, 0), -8),
                       ((None, N * es * g, None, N * es, None, ks, None, N * 16 - 1, 0), -5),
                       ((None, N * es * g, None, N * es, None, ks, None, N * 16, 3), -12)]:
        full = (*args, n, m, g, es, N, None, 0, st)
        assert lib.FastFss_hip_dcfEvalNoAlloc(*full) == lib.FastFss_cuda_dcfEval(*full) == code
    N = MAX_ELEMS + 1
    ks = key_size(dev, "dcfGetKeyDataSize", n, m, g, es, N)
    assert lib.FastFss_hip_dcfEvalNoAlloc(None, N * es * g, None, N * es, None, ks, None, N * 16, 0, n, m, g, es, N,
                                          None, 0, st) == -1
    assert lib.FastFss_hip_dcfEvalNoAlloc(None, 0, None, 0, None, 0, None, 0, 2, n, m, g, es, 0, None, 0, st) == -12
