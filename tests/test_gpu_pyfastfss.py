"""pyFastFss on GPU tensors: the reference binding surface (same names,
arguments, reshaping) driven with torch tensors, every key and share compared
byte for byte with the oracle on the same inputs, and DCF shares reconstructed
to the comparison value."""
import numpy as np
import pytest

import fss_testlib as F

pytestmark = pytest.mark.gpu

O = F.OracleBackend()
UDT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


@pytest.fixture(scope="module")
def env():
    import torch

    from fastfunctionsecretsharing_amd import pyFastFss as P

    torch.cuda.set_device(0)
    return torch, P


def _np(t, es):
    return t.detach().cpu().contiguous().numpy().view(UDT[es]).reshape(-1)


def _tt(torch, a, dtype, shape=None):
    t = torch.from_numpy(np.ascontiguousarray(a).view(np.dtype(str(dtype).replace("torch.", "")))).to("cuda:0")
    return t.reshape(shape) if shape is not None else t


def _inputs(torch, P, rng, shape, n, es):
    N = int(np.prod(shape))
    dt = P.dtype(8 * es)
    alpha = F.rand_elems(rng, N, n, es)
    x = F.rand_elems(rng, N, n, es)
    s0, s1 = F.rand_seeds(rng, N), F.rand_seeds(rng, N)
    return N, dt, alpha, x, s0, s1


@pytest.mark.parametrize("n,m,es,g", [(32, 32, 4, 1), (16, 8, 2, 2), (64, 64, 8, 1), (8, 8, 1, 3)])
def test_dcf(env, n, m, es, g):
    torch, P = env
    rng = np.random.default_rng(n * 7 + g)
    shape = (3, 41)
    N, dt, alpha, x, s0, s1 = _inputs(torch, P, rng, shape, n, es)
    beta = F.rand_elems(rng, N * g, m, es)
    key = torch.empty(0, dtype=torch.uint8, device="cuda:0")
    P.dcf_key_gen(key, _tt(torch, alpha, dt, shape), _tt(torch, beta, dt), _tt(torch, s0, torch.uint8),
                  _tt(torch, s1, torch.uint8), n, m, g)
    okey = F.dcf_keygen(O, alpha, beta, s0, s1, n, m, g, es)
    assert key.numel() == P.dcf_get_key_data_size(n, m, g, es, N)
    assert np.array_equal(key.cpu().numpy(), okey)
    xm = _tt(torch, x, dt, shape)
    outs = []
    for party, sd in ((0, s0), (1, s1)):
        out = torch.empty(0, dtype=dt, device="cuda:0")
        r = P.dcf_eval(out, xm, key, _tt(torch, sd, torch.uint8), party, n, m, g)
        assert r is out and tuple(out.shape) == shape + (g,)
        assert np.array_equal(_np(out, es), F.dcf_eval(O, x, okey, sd, party, n, m, g, es))
        outs.append(out)
    rec = (_np(outs[0], es).astype(np.uint64) + _np(outs[1], es)) & np.uint64(F.mask(m))
    # shares reconstruct to beta * [x < alpha] on the evaluated input
    want = np.where((x < alpha)[:, None], beta.reshape(N, g).astype(np.uint64), 0)
    assert np.array_equal(rec.reshape(N, g), want)


@pytest.mark.parametrize("n,m,es", [(20, 16, 4), (12, 8, 2)])
def test_dpf_family(env, n, m, es):
    torch, P = env
    rng = np.random.default_rng(n + m)
    shape = (2, 17)
    N, dt, alpha, x, s0, s1 = _inputs(torch, P, rng, shape, n, es)
    x[::3] = alpha[::3]
    beta = F.rand_elems(rng, N, m, es)
    key = torch.empty(0, dtype=torch.uint8, device="cuda:0")
    P.dpf_key_gen(key, _tt(torch, alpha, dt, shape), _tt(torch, beta, dt), _tt(torch, s0, torch.uint8),
                  _tt(torch, s1, torch.uint8), n, m, 1)
    okey = F.dpf_keygen(O, alpha, beta, s0, s1, n, m, 1, es)
    assert np.array_equal(key.cpu().numpy(), okey)
    xm = _tt(torch, x, dt, shape)
    points = F.rand_elems(rng, 5, n, es)
    points[0] = 0
    for party, sd in ((0, s0), (1, s1)):
        sdt = _tt(torch, sd, torch.uint8)
        out = torch.empty(0, dtype=dt, device="cuda:0")
        P.dpf_eval(out, xm, key, sdt, party, n, m, 1)
        assert tuple(out.shape) == shape
        assert np.array_equal(_np(out, es), F.dpf_eval(O, x, okey, sd, party, n, m, 1, es))
        mo = torch.empty(0, dtype=dt, device="cuda:0")
        P.dpf_multi_eval(mo, xm, key, sdt, party, _tt(torch, points, dt), n, m, 1)
        assert tuple(mo.shape) == shape + (5,)
        assert np.array_equal(_np(mo, es), F.dpf_eval_multi(O, x, okey, sd, party, points, n, m, 1, es))
    if n <= 12:
        for party, sd in ((0, s0), (1, s1)):
            ao = torch.empty(0, dtype=dt, device="cuda:0")
            P.dpf_eval_all(ao, xm, key, _tt(torch, sd, torch.uint8), party, n, m, 1)
            assert tuple(ao.shape) == shape + (1 << n,)
            assert np.array_equal(_np(ao, es), F.dpf_eval_all(O, x, okey, sd, party, n, m, 1, es))


@pytest.mark.parametrize("n,es", [(16, 2), (32, 4)])
def test_grotto_family(env, n, es):
    torch, P = env
    rng = np.random.default_rng(n * 3)
    shape = (5, 13)
    N, dt, alpha, x, s0, s1 = _inputs(torch, P, rng, shape, n, es)
    x[::4] = alpha[::4]
    key = torch.empty(0, dtype=torch.uint8, device="cuda:0")
    P.grotto_key_gen(key, _tt(torch, alpha, dt, shape), _tt(torch, s0, torch.uint8), _tt(torch, s1, torch.uint8), n)
    okey = F.grotto_keygen(O, alpha, s0, s1, n, es)
    assert np.array_equal(key.cpu().numpy(), okey)
    xm = _tt(torch, x, dt, shape)
    step = 1 << (n - 3)
    left = np.array([i * step for i in range(8)], dtype=UDT[es])
    right = np.array([(i + 1) * step - 1 for i in range(8)], dtype=UDT[es])
    lut = F.rand_elems(rng, 8 * 2, n, es)
    for party, sd in ((0, s0), (1, s1)):
        sdt = _tt(torch, sd, torch.uint8)
        for eb in (False, True):
            o = torch.empty(0, dtype=dt, device="cuda:0")
            P.grotto_eval(o, xm, key, sdt, eb, party, n)
            assert tuple(o.shape) == shape
            assert np.array_equal(_np(o, es), F.grotto_eval(O, x, okey, sd, party, n, es, eb))
        o = torch.empty(0, dtype=dt, device="cuda:0")
        P.grotto_eq_eval(o, xm, key, sdt, party, n)
        assert np.array_equal(_np(o, es), F.grotto_eq_eval(O, x, okey, sd, party, n, es))
        o = torch.empty(0, dtype=dt, device="cuda:0")
        P.grotto_mic_eval(o, xm, key, sdt, party, _tt(torch, left, dt), _tt(torch, right, dt), n)
        assert tuple(o.shape) == shape + (8,)
        assert np.array_equal(_np(o, es), F.grotto_mic_eval(O, x, okey, sd, party, left, right, n, es))
        oe = torch.empty(0, dtype=dt, device="cuda:0")
        ot = torch.empty(0, dtype=dt, device="cuda:0")
        P.grotto_interval_lut_eval(oe, ot, xm, key, sdt, party, _tt(torch, left, dt), _tt(torch, right, dt),
                                   _tt(torch, lut, dt), n, n)
        assert tuple(oe.shape) == shape and tuple(ot.shape) == shape + (2,)
        we, wt = F.grotto_lut_eval(O, x, okey, sd, party, left, right, lut, n, n, es)
        assert np.array_equal(_np(oe, es), we) and np.array_equal(_np(ot, es), wt)


def test_dcf_mic(env):
    torch, P = env
    n = m = 32
    es = 4
    rng = np.random.default_rng(5)
    shape = (4, 25)
    N, dt, alpha, x, s0, s1 = _inputs(torch, P, rng, shape, n, es)
    left = np.array([(i * 0x1000_0000) & 0xFFFFFFFF for i in range(11)], dtype=np.uint32)
    right = np.array([(i * 0x1000_0000 + 0x0FFF_FFFF) & 0xFFFFFFFF for i in range(11)], dtype=np.uint32)
    key = torch.empty(0, dtype=torch.uint8, device="cuda:0")
    z = torch.empty(0, dtype=dt, device="cuda:0")
    P.dcf_mic_key_gen(key, z, _tt(torch, alpha, dt, shape), _tt(torch, s0, torch.uint8),
                      _tt(torch, s1, torch.uint8), _tt(torch, left, dt), _tt(torch, right, dt), n, m)
    assert tuple(z.shape) == shape + (11,)
    okey, oz = F.mic_keygen(O, alpha, s0, s1, left, right, n, m, es)
    assert np.array_equal(key.cpu().numpy(), okey) and np.array_equal(_np(z, es), oz)
    z0 = torch.randint_like(z, -(2**31), 2**31 - 1)
    z1 = z - z0
    xm = _tt(torch, x, dt, shape)
    for party, sd, zs in ((0, s0, z0), (1, s1, z1)):
        o = torch.empty(0, dtype=dt, device="cuda:0")
        P.dcf_mic_eval(o, xm, key, zs, _tt(torch, sd, torch.uint8), party, _tt(torch, left, dt),
                       _tt(torch, right, dt), n, m)
        assert tuple(o.shape) == shape + (11,)
        assert np.array_equal(_np(o, es), F.mic_eval(O, x, okey, _np(zs, es), sd, party, left, right, n, m, es))


@pytest.mark.parametrize("degree", [1, 2])
def test_spline(env, degree):
    torch, P = env
    n = m = 16
    es = 2
    rng = np.random.default_rng(degree)
    shape = (6, 9)
    N, dt, alpha, x, s0, s1 = _inputs(torch, P, rng, shape, n, es)
    I = 4
    left = np.array([i << 14 for i in range(I)], dtype=np.uint16)
    right = np.array([(i << 14) + (1 << 14) - 1 for i in range(I)], dtype=np.uint16)
    coeff = F.rand_elems(rng, I * (degree + 1), m, es)
    key = torch.empty(0, dtype=torch.uint8, device="cuda:0")
    e = torch.empty(0, dtype=dt, device="cuda:0")
    b = torch.empty(0, dtype=dt, device="cuda:0")
    P.spline_key_gen(key, e, b, _tt(torch, alpha, dt, shape), _tt(torch, s0, torch.uint8),
                     _tt(torch, s1, torch.uint8), _tt(torch, coeff, dt), degree, _tt(torch, left, dt),
                     _tt(torch, right, dt), I, n, m)
    assert tuple(e.shape) == shape + (I * (degree + 1),) and e.shape == b.shape
    okey, oe, ob = F.spline_keygen(O, alpha, s0, s1, coeff, degree, left, right, n, m, es)
    assert np.array_equal(key.cpu().numpy(), okey)
    assert np.array_equal(_np(e, es), oe) and np.array_equal(_np(b, es), ob)
    e0 = torch.randint_like(e, -(2**15), 2**15 - 1)
    b0 = torch.randint_like(b, -(2**15), 2**15 - 1)
    xm = _tt(torch, x, dt, shape)
    for party, sd, ee, bb in ((0, s0, e0, b0), (1, s1, e - e0, b - b0)):
        o = torch.empty(0, dtype=dt, device="cuda:0")
        P.spline_eval(o, xm, key, ee, bb, _tt(torch, sd, torch.uint8), party, _tt(torch, left, dt),
                      _tt(torch, right, dt), I, degree, n, m)
        assert tuple(o.shape) == shape
        want = F.spline_eval(O, x, okey, _np(ee, es), _np(bb, es), sd, party, left, right, degree, n, m, es)
        assert np.array_equal(_np(o, es), want)


def test_ottt(env):
    torch, P = env
    n, es = 8, 2
    rng = np.random.default_rng(11)
    shape = (7, 5)
    N, dt, alpha, x, s0, s1 = _inputs(torch, P, rng, shape, n, es)
    ks = P.ottt_get_key_data_size(n, N)
    key0 = rng.integers(0, 256, size=ks, dtype=np.uint8)
    key = _tt(torch, key0, torch.uint8)
    P.ottt_key_gen(key, _tt(torch, alpha, dt, shape), n)
    assert key.numel() == ks
    okey = F.ottt_keygen(O, key0, alpha, n, es)
    assert np.array_equal(key.cpu().numpy(), okey)
    lut = F.rand_elems(rng, 2 << n, 16, es)
    xm = _tt(torch, x, dt, shape)
    for party, kk, ok in ((0, _tt(torch, key0, torch.uint8), key0), (1, key, okey)):
        oe = torch.empty(0, dtype=dt, device="cuda:0")
        ot = torch.empty(0, dtype=dt, device="cuda:0")
        P.ottt_lut_eval(oe, ot, xm, kk, party, _tt(torch, lut, dt), n, 16)
        assert tuple(oe.shape) == shape and tuple(ot.shape) == shape + (2,)
        we, wt = F.ottt_eval(O, x, ok, party, lut, n, es)
        assert np.array_equal(_np(oe, es), we) and np.array_equal(_np(ot, es), wt)


def test_prng(env):
    torch, P = env
    g = P.Prng(torch.device("cuda:0"))
    assert g.device() == torch.device("cuda:0")
    seed, ctr = bytes(range(16)), bytes(16)
    g.set_current_seed(seed, ctr)
    a = torch.empty(1000, dtype=torch.int32, device="cuda:0")
    g.rand_(a, 20)
    s2, c2 = g.get_current_seed()
    assert s2 == seed and c2 != ctr
    o = F.oracle()
    h = o.oracle_prngInit()
    o.oracle_prngSetCurrentSeed(h, seed, ctr)
    want = np.zeros(1000, dtype=np.uint32)
    assert o.oracle_prngGen(h, want.ctypes.data, 20, 4, 1000) == 0
    o.oracle_prngRelease(h)
    assert np.array_equal(_np(a, 4), want)
    with pytest.raises(ValueError):
        g.rand_(torch.empty(4, dtype=torch.int8, device="cuda:0"), 9)


def test_error_code_surface(env):
    torch, P = env
    x = torch.zeros(8, dtype=torch.int32, device="cuda:0")
    key = torch.zeros(3, dtype=torch.uint8, device="cuda:0")
    seed = torch.zeros(128, dtype=torch.uint8, device="cuda:0")
    out = torch.empty(0, dtype=torch.int32, device="cuda:0")
    with pytest.raises(RuntimeError, match=r"FastFss_cuda_dcfEval failed : -?\d+"):
        P.dcf_eval(out, x, key, seed, 0, 32, 32, 1)
