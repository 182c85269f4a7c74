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
                         mic_eval, mic_keygen, rand_elems, rand_seeds, spline_closed_form, spline_eval,
                         spline_keygen, sub_key_rows)

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


def _spline_setup(rng, N, n, m, es, I, d, gap=True):
    """Contiguous sorted intervals covering [0, 2^n) with one gap (interval 3 starts one later),
    device keys, and beta/e split into random additive shares."""
    M = mask(n)
    cuts = sorted(set(int(v) for v in rand_elems(rng, 4 * I, n, es)))[:I - 1]
    right = np.array(cuts + [M], dtype=DT[es])
    left = np.array([0] + [c + 1 for c in cuts], dtype=DT[es])
    if gap:
        left[3] = (int(left[3]) + 1) & M
    coeff = rand_elems(rng, I * (d + 1), m, es)
    alpha, masked, s0, s1 = _inputs(rng, N, n, es)
    masked[:I] = ((left.astype(np.uint64) + alpha[:I]) & np.uint64(M)).astype(DT[es])
    masked[I] = ((int(left[3]) - 1 + int(alpha[I])) & M)  # the gap's point
    kd, ed, bd = spline_keygen(dev_be(), alpha, s0, s1, coeff, d, left, right, n, m, es)
    e, b = dev_be().get(ed), dev_be().get(bd)
    W = np.uint64(mask(8 * es))
    b0, e0 = rand_elems(rng, len(b), 8 * es, es), rand_elems(rng, len(e), 8 * es, es)
    b1 = ((b.astype(np.uint64) - b0) & W).astype(DT[es])
    e1 = ((e.astype(np.uint64) - e0) & W).astype(DT[es])
    return left, right, coeff, alpha, masked, (s0, s1), kd, (e0, e1), (b0, b1)


_DEV = []


def dev_be():
    if not _DEV:
        _DEV.append(DeviceBackend())
    return _DEV[0]


@pytest.mark.parametrize("N", [16385, 40000, 70000, 131073])
def test_spline_split_shapes(dev, N):
    """Spline calls too small to fill the GPU split an element's intervals over 2..16 lanes (blocks of
    linked intervals, halved further when the grid is capped). At these sizes, with 16 intervals and
    one gap (n = 24, m = 32, degree 2): keys of a sample equal the oracle's, both parties' shares of
    the sample equal the oracle's, and every element reconstructs to its closed form."""
    n, m, es, I, d = 24, 32, 4, 16, 2
    rng = np.random.default_rng(N + 3)
    left, right, coeff, alpha, masked, seeds, kd, es_, bs = _spline_setup(rng, N, n, m, es, I, d)
    g = I * (d + 1)
    idx = np.unique(np.concatenate([rng.choice(N, 200, replace=False), np.arange(I + 1), [N - 1]]))
    ko = spline_keygen(OR, alpha[idx], seeds[0].reshape(N, 16)[idx].ravel(), seeds[1].reshape(N, 16)[idx].ravel(),
                       coeff, d, left, right, n, m, es)[0]
    assert np.array_equal(sub_key_rows(dev.get(kd), idx, 16 * n, (n * g + 2 + g) * es, N), ko)
    ys = []
    for party in (0, 1):
        y = dev.get(spline_eval(dev, masked, kd, es_[party], bs[party], seeds[party], party, left, right, d, n, m, es))
        yo = spline_eval(OR, masked[idx], ko, es_[party].reshape(N, g)[idx].ravel(),
                         bs[party].reshape(N, g)[idx].ravel(), seeds[party].reshape(N, 16)[idx].ravel(), party, left,
                         right, d, n, m, es)
        assert np.array_equal(y[idx], yo), party
        ys.append(y)
    rec = (ys[0].astype(np.uint64) + ys[1]) & np.uint64(mask(m))
    assert np.array_equal(rec, spline_closed_form(masked, alpha, left, right, coeff, d, n, m))


@pytest.mark.parametrize("n,m,es,I,d,N", [(24, 32, 4, 16, 2, 17), (24, 32, 4, 16, 2, 1000), (24, 32, 4, 16, 2, 4096),
                                          (32, 32, 4, 4, 2, 8192), (24, 32, 4, 16, 2, 8192), (40, 32, 8, 5, 3, 1000),
                                          (20, 30, 4, 6, 1, 300)])
def test_spline_small_call_lanes(dev, n, m, es, I, d, N):
    """Small spline calls with 4-byte slots walk each DCF point on eight lanes (column-parallel AES)
    with 8-, 4- and 2-way interval splits at these sizes (16 intervals at 8192 elements would be 8
    per item, so that call takes the split kernel); degree 2 makes linked walks that touch three
    conversion windows, 64-bit elements carry 4-byte outputs, and interval 3 starts after a gap.
    Keys of a sample equal the oracle's, both parties' shares of the sample equal the oracle's, and
    every element reconstructs to its closed form."""
    rng = np.random.default_rng(N + 7 * n + d)
    left, right, coeff, alpha, masked, seeds, kd, es_, bs = _spline_setup(rng, N, n, m, es, I, d)
    g = I * (d + 1)
    idx = np.unique(np.concatenate([rng.choice(N, min(N, 200), replace=False), np.arange(I + 1), [N - 1]]))
    ko = spline_keygen(OR, alpha[idx], seeds[0].reshape(N, 16)[idx].ravel(), seeds[1].reshape(N, 16)[idx].ravel(),
                       coeff, d, left, right, n, m, es)[0]
    assert np.array_equal(sub_key_rows(dev.get(kd), idx, 16 * n, (n * g + 2 + g) * es, N), ko)
    ys = []
    for party in (0, 1):
        y = dev.get(spline_eval(dev, masked, kd, es_[party], bs[party], seeds[party], party, left, right, d, n, m, es))
        yo = spline_eval(OR, masked[idx], ko, es_[party].reshape(N, g)[idx].ravel(),
                         bs[party].reshape(N, g)[idx].ravel(), seeds[party].reshape(N, 16)[idx].ravel(), party, left,
                         right, d, n, m, es)
        assert np.array_equal(y[idx], yo), party
        ys.append(y)
    rec = (ys[0].astype(np.uint64) + ys[1]) & np.uint64(mask(m))
    assert np.array_equal(rec, spline_closed_form(masked, alpha, left, right, coeff, d, n, m))


@pytest.mark.parametrize("n,m,es", [(24, 32, 4), (12, 10, 2)])
@pytest.mark.parametrize("N", [1000, 300000])
def test_spline_reconstruction_in_ne_out(dev, n, m, es, N):
    """bitWidthIn != bitWidthOut on the split kernel (1000 elements) and the large-call kernel
    (300000): every element's two shares reconstruct to f_q(X - alpha) mod 2^m at the integer
    X - alpha, 0 in the gap, computed from the mathematics alone."""
    I, d = 8, 2
    rng = np.random.default_rng(N + n)
    left, right, coeff, alpha, masked, seeds, kd, es_, bs = _spline_setup(rng, N, n, m, es, I, d)
    ys = [dev.get(spline_eval(dev, masked, kd, es_[p], bs[p], seeds[p], p, left, right, d, n, m, es)) for p in (0, 1)]
    rec = (ys[0].astype(np.uint64) + ys[1]) & np.uint64(mask(m))
    want = spline_closed_form(masked, alpha, left, right, coeff, d, n, m)
    assert np.array_equal(rec, want)
    assert (masked < alpha).any() and (want == 0).any()
