"""Per-call launches across the workgroup-size boundaries.

A call is spread over all CUs in workgroups of 64..1024 lanes (whole waves)
until it fills the GPU; the static-split kernels serve calls below 2^16
elements and the work-queue kernels (small-workgroup and persistent forms)
the rest.  Each size below lands on a different shape: one 64-lane
workgroup, two, 128-lane workgroups on fewer CUs than the GPU has, 256-lane
work-queue workgroups, and the persistent 1024-lane form.  Keys and shares
must equal the oracle's, bit for bit.
"""
import numpy as np
import pytest

from fss_testlib import (DT, DeviceBackend, OracleBackend, dcf_eval, dcf_keygen, dpf_eval, dpf_keygen, mask,
                         mic_eval, mic_keygen, rand_elems, rand_seeds)

pytestmark = pytest.mark.gpu

OR = OracleBackend()
SIZES = [64, 65, 16385, 65536, 262145]


@pytest.fixture(scope="module")
def dev():
    return DeviceBackend()


def _inputs(rng, N, n, es):
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    x[::7] = alpha[::7]
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    return alpha, masked, rand_seeds(rng, N), rand_seeds(rng, N)


@pytest.mark.parametrize("N", SIZES)
def test_dcf_shapes(dev, N):
    n, m, es = 20, 20, 4
    rng = np.random.default_rng(N)
    alpha, masked, s0, s1 = _inputs(rng, N, n, es)
    beta = rand_elems(rng, N, m, es)
    ko = dcf_keygen(OR, alpha, beta, s0, s1, n, m, 1, es)
    kd = dcf_keygen(dev, alpha, beta, s0, s1, n, m, 1, es)
    assert np.array_equal(dev.get(kd), ko)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dev.get(dcf_eval(dev, masked, kd, sd, party, n, m, 1, es)),
                              dcf_eval(OR, masked, ko, sd, party, n, m, 1, es)), party


@pytest.mark.parametrize("N", SIZES)
def test_dpf_shapes(dev, N):
    n, m, es = 20, 24, 4
    rng = np.random.default_rng(N + 1)
    alpha, masked, s0, s1 = _inputs(rng, N, n, es)
    beta = rand_elems(rng, N, m, es)
    ko = dpf_keygen(OR, alpha, beta, s0, s1, n, m, 1, es)
    kd = dpf_keygen(dev, alpha, beta, s0, s1, n, m, 1, es)
    assert np.array_equal(dev.get(kd), ko)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dev.get(dpf_eval(dev, masked, kd, sd, party, n, m, 1, es)),
                              dpf_eval(OR, masked, ko, sd, party, n, m, 1, es)), party


@pytest.mark.parametrize("N", [64, 1000, 16385])
def test_mic_shapes(dev, N):
    """MIC splits each element's intervals over lanes, so these sizes give
    64- to 1024-lane workgroups (small-workgroup and persistent kernels)."""
    n, m, es = 16, 16, 4
    rng = np.random.default_rng(N + 2)
    alpha, masked, s0, s1 = _inputs(rng, N, n, es)
    left = np.array([0, 100, 5000, 40000], dtype=DT[es])
    right = np.array([99, 4999, 39999, 65535], dtype=DT[es])
    ko, zo = mic_keygen(OR, alpha, s0, s1, left, right, n, m, es)
    kd, zd = mic_keygen(dev, alpha, s0, s1, left, right, n, m, es)
    assert np.array_equal(dev.get(kd), ko)
    assert np.array_equal(dev.get(zd), zo)
    for party, sd in ((0, s0), (1, s1)):
        assert np.array_equal(dev.get(mic_eval(dev, masked, kd, zd, sd, party, left, right, n, m, es)),
                              mic_eval(OR, masked, ko, zo, sd, party, left, right, n, m, es)), party
