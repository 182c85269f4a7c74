"""Elements of 16, 24 and 32 bytes (the reference's wideint::uint<2>, <3>, <4>,
src/impl/def.h:37-44) through every FastFss_cuda_* entry point.

Three things pin them:
* For bitWidthIn and bitWidthOut <= 64, a wide key and its shares equal the
  8-byte path's (itself pinned against the oracle) in their low 64 bits: the
  seed corrections are the same bytes, the value conversion the same bytes,
  and arithmetic mod 2^64 is the low part of arithmetic mod 2^(8 es).  The
  control-bit words hold no bits above n.
* At bitWidthIn = bitWidthOut = 8 es the two parties' shares reconstruct, in
  Python integers, to each primitive's function over the full width.
* The value conversion of outputs wider than 16 bytes (an AES refill keyed by
  the value block, src/impl/dcf.h:99-129), which reconstruction cancels, is
  checked on a depth-1 DCF key against AES from the system's libcrypto.
"""
import numpy as np
import pytest

from fss_testlib import (Aes, DeviceBackend, dcf_eval, dcf_keygen, dpf_eval, dpf_keygen, grotto_eq_eval,
                         grotto_eval, grotto_keygen, key_size, mic_eval, mic_keygen, ottt_eval, ottt_keygen,
                         rand_seeds, run, spline_eval, spline_keygen)

pytestmark = pytest.mark.gpu

WIDE = (16, 24, 32)


@pytest.fixture(scope="module")
def dev():
    return DeviceBackend()


def enc(vals, es):
    """Python ints -> little-endian bytes, es per value (mod 2^(8 es))."""
    return np.frombuffer(b"".join((int(v) % (1 << (8 * es))).to_bytes(es, "little") for v in vals),
                         np.uint8).copy()


def dec(buf, es):
    b = np.asarray(buf, dtype=np.uint8).tobytes()
    return [int.from_bytes(b[i:i + es], "little") for i in range(0, len(b), es)]


def widen(a64, es):
    """uint64 values -> the same values as es-byte elements."""
    return enc([int(v) for v in np.asarray(a64, dtype=np.uint64)], es)


def low64(buf, es):
    return np.array([v & (2**64 - 1) for v in dec(buf, es)], dtype=np.uint64)


def rbits(rng, count, bits):
    return [int.from_bytes(rng.bytes((bits + 7) // 8), "little") % (1 << bits) for _ in range(count)]


def dget(dev, t):
    return dev.get(t).view(np.uint8)


def wcall(dev, name, *args):
    run(dev, name, *args)


# ------------------------------------------------------------------- DCF ----

def _dcf_wide(dev, alpha_b, beta_b, s0, s1, n, m, g, es, N):
    ks = key_size(dev, "dcfGetKeyDataSize", n, m, g, es, N)
    key = dev.empty(ks)
    a, d0, d1 = dev.put(alpha_b), dev.put(s0), dev.put(s1)
    b = dev.put(beta_b) if beta_b is not None else None
    wcall(dev, "dcfKeyGen", key, ks, a, N * es, b, 0 if b is None else N * es * g, d0, N * 16, d1, N * 16, n, m, g,
          es, N)
    return key


def _dcf_wide_eval(dev, x_b, key, seed, party, n, m, g, es, N):
    out = dev.empty(N * es * g)
    wcall(dev, "dcfEval", out, N * es * g, dev.put(x_b), N * es, key, key.numel(), dev.put(seed), N * 16, party, n, m,
          g, es, N, None, 0)
    return out


def _tail_low64_equal(kw, kn, head, nelem_w, nelem_n, es):
    """Key bytes after `head` bytes: wide elements' low 64 bits equal the 8-byte elements."""
    assert np.array_equal(kw[:head], kn[:head])
    assert len(kw) - head == nelem_w * es and len(kn) - head == nelem_n * 8
    assert np.array_equal(low64(kw[head:], es), kn[head:].view(np.uint64))


@pytest.mark.parametrize("es", WIDE)
@pytest.mark.parametrize("g", [1, 3])
def test_dcf_low64_matches_8byte(dev, es, g):
    rng = np.random.default_rng(es * 10 + g)
    N, n, m = 300, 48, 40
    alpha = rng.integers(0, 2**n, N, dtype=np.uint64)
    beta = rng.integers(0, 2**63, N * g, dtype=np.uint64) if g > 1 else None
    x = rng.integers(0, 2**n, N, dtype=np.uint64)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kn = dev.get(dcf_keygen(dev, alpha, beta, s0, s1, n, m, g, 8)).view(np.uint8)
    kw = dget(dev, _dcf_wide(dev, widen(alpha, es), None if beta is None else widen(beta, es), s0, s1, n, m, g, es,
                             N))
    _tail_low64_equal(kw, kn, N * 16 * n, N * (n * g + 2 + g), N * (n * g + 2 + g), es)
    tails = np.array(dec(kw[N * 16 * n:], es), dtype=object).reshape(N, n * g + 2 + g)
    assert all(int(v) >> 64 == 0 for v in tails[:, n * g]) and all(int(v) >> 64 == 0 for v in tails[:, n * g + 1])
    kd = dev.put(kw)
    for party, seed in ((0, s0), (1, s1)):
        yn = dev.get(dcf_eval(dev, x, dev.put(kn), seed, party, n, m, g, 8))
        yw = dget(dev, _dcf_wide_eval(dev, widen(x, es), kd, seed, party, n, m, g, es, N))
        assert np.array_equal(low64(yw, es), yn.view(np.uint64)), party


@pytest.mark.parametrize("es", WIDE)
@pytest.mark.parametrize("m_minus", [0, 56])
def test_dcf_full_width_reconstruction(dev, es, m_minus):
    rng = np.random.default_rng(es + m_minus)
    N, g = 96, 2
    n = 8 * es
    m = 8 * es - m_minus
    alpha = rbits(rng, N, n)
    masked = rbits(rng, N, n)  # the evaluation points
    masked[:8] = alpha[:8]
    masked[8:16] = [(a - 1) % (1 << n) for a in alpha[8:16]]
    masked[16:24] = [(a + 1) % (1 << n) for a in alpha[16:24]]
    beta = rbits(rng, N * g, 8 * es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = _dcf_wide(dev, enc(alpha, es), enc(beta, es), s0, s1, n, m, g, es, N)
    ys = [dec(dget(dev, _dcf_wide_eval(dev, enc(masked, es), key, sd, p, n, m, g, es, N)), es)
          for p, sd in ((0, s0), (1, s1))]
    M = (1 << (8 * es)) - 1
    for i in range(N):
        for j in range(g):
            want = beta[i * g + j] if masked[i] < alpha[i] else 0
            assert (ys[0][i * g + j] + ys[1][i * g + j]) & M == want, (i, j)


def _aes(key16: bytes, lo: int, hi: int) -> bytes:
    return Aes.encrypt(key16, lo.to_bytes(8, "little") + hi.to_bytes(8, "little"))


def _conv_g1(v: bytes, nbytes: int) -> int:
    """groupSize-1 conversion of a 16-byte value block into nbytes > 16 bytes: the buffer is refilled
    with AES_v(0 || element index) whenever the byte index wraps to 0, and the index stays 0."""
    out = b""
    for i in range(nbytes):
        if i % 16 == 0:
            buf = _aes(v, 0, 0)
        out += buf[i % 16:i % 16 + 1]
    return int.from_bytes(out, "little")


@pytest.mark.skipif(not Aes.available(), reason="the system's libcrypto is not loadable")
@pytest.mark.parametrize("m", [136, 200, 256])
def test_dcf_wide_conversion_closed_form(dev, m):
    """Depth-1 DCF keys with bitWidthOut > 128 (32-byte elements) against a closed form of the first
    level built on libcrypto AES (src/impl/dcf.h:187-368 at n = 1)."""
    es, n, g, N = 32, 1, 1, 16
    rng = np.random.default_rng(m)
    alpha = rbits(rng, N, 1)
    beta = rbits(rng, N, 8 * es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dget(dev, _dcf_wide(dev, enc(alpha, es), enc(beta, es), s0, s1, n, m, g, es, N))
    M = (1 << 256) - 1
    nb = (m + 7) // 8
    for i in range(N):
        seeds = []
        for sd in (s0, s1):
            b = bytearray(sd[16 * i:16 * i + 16].tobytes())
            b[0] &= 0xFE
            seeds.append(bytes(b))
        blocks = [[_aes(sd, 2 * q, 2 * q + 1) for q in range(4)] for sd in seeds]  # sL vL sR vR
        tb = [[blk[0][0] & 1, blk[2][0] & 1] for blk in blocks]
        sd_ = [[bytes([blk[k][0] & 0xFE]) + blk[k][1:] for k in (0, 2)] for blk in blocks]
        keep = alpha[i] & 1
        lose = 1 - keep
        sCW = bytes(a ^ b for a, b in zip(sd_[0][lose], sd_[1][lose]))
        sign1 = -1  # t1 = 1 at the root
        vl = [blocks[p][1 + 2 * lose] for p in (0, 1)]
        vk = [blocks[p][1 + 2 * keep] for p in (0, 1)]
        vcw = (sign1 * (_conv_g1(vl[1], nb) - _conv_g1(vl[0], nb))) & M
        if lose == 0:
            vcw = (vcw + sign1 * beta[i]) & M
        valpha = (-_conv_g1(vk[1], nb) + _conv_g1(vk[0], nb) + sign1 * vcw) & M
        tLCW = tb[0][0] ^ tb[1][0] ^ keep ^ 1
        tRCW = tb[0][1] ^ tb[1][1] ^ keep
        tkcw = tRCW if keep else tLCW
        ns0 = sd_[0][keep]
        ns1 = bytes(a ^ b for a, b in zip(sd_[1][keep], sCW))
        t1 = tb[1][keep] ^ tkcw
        last = ((-valpha + _conv_g1(ns1, nb) - _conv_g1(ns0, nb)) * (-1 if t1 else 1)) & M
        assert key[16 * i:16 * i + 16].tobytes() == sCW, i
        tail = dec(key[N * 16 + i * es * 4:N * 16 + (i + 1) * es * 4], es)
        assert tail == [vcw, tLCW, tRCW, last], i


# ------------------------------------------------------------------- DPF ----

def _dpf_wide(dev, alpha_b, beta_b, s0, s1, n, m, g, es, N):
    ks = key_size(dev, "dpfGetKeyDataSize", n, m, g, es, N)
    key = dev.empty(ks)
    b = dev.put(beta_b) if beta_b is not None else None
    wcall(dev, "dpfKeyGen", key, ks, dev.put(alpha_b), N * es, b, 0 if b is None else N * es * g, dev.put(s0), N * 16,
          dev.put(s1), N * 16, n, m, g, es, N)
    return key


@pytest.mark.parametrize("es", WIDE)
def test_dpf_low64_and_full_width(dev, es):
    rng = np.random.default_rng(100 + es)
    N, n, m, g = 200, 40, 48, 2
    alpha = rng.integers(0, 2**n, N, dtype=np.uint64)
    beta = rng.integers(0, 2**63, N * g, dtype=np.uint64)
    x = rng.integers(0, 2**n, N, dtype=np.uint64)
    x[::4] = alpha[::4]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kn = dev.get(dpf_keygen(dev, alpha, beta, s0, s1, n, m, g, 8)).view(np.uint8)
    kw = dget(dev, _dpf_wide(dev, widen(alpha, es), widen(beta, es), s0, s1, n, m, g, es, N))
    _tail_low64_equal(kw, kn, N * 16 * n, N * (2 + g), N * (2 + g), es)
    kd = dev.put(kw)
    for party, seed in ((0, s0), (1, s1)):
        yn = dev.get(dpf_eval(dev, x, dev.put(kn), seed, party, n, m, g, 8))
        out = dev.empty(N * es * g)
        wcall(dev, "dpfEval", out, N * es * g, dev.put(widen(x, es)), N * es, kd, kd.numel(), dev.put(seed), N * 16,
              party, n, m, g, es, N, None, 0)
        assert np.array_equal(low64(dget(dev, out), es), yn.view(np.uint64)), party
    # full width: n = m = 8 es, shares reconstruct to beta [x == alpha]
    n = m = 8 * es
    N = 64
    alpha = rbits(rng, N, n)
    beta = rbits(rng, N, 8 * es)
    x = rbits(rng, N, n)
    x[::3] = alpha[::3]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = _dpf_wide(dev, enc(alpha, es), enc(beta, es), s0, s1, n, m, 1, es, N)
    ys = []
    for party, seed in ((0, s0), (1, s1)):
        out = dev.empty(N * es)
        wcall(dev, "dpfEval", out, N * es, dev.put(enc(x, es)), N * es, key, key.numel(), dev.put(seed), N * 16, party,
              n, m, 1, es, N, None, 0)
        ys.append(dec(dget(dev, out), es))
    M = (1 << (8 * es)) - 1
    for i in range(N):
        assert (ys[0][i] + ys[1][i]) & M == (beta[i] if x[i] == alpha[i] else 0), i


@pytest.mark.parametrize("es", [16, 32])
def test_dpf_eval_all_and_multi(dev, es):
    """evalAll at n = 6 and evalMulti equal the per-point eval at the points they stand for."""
    rng = np.random.default_rng(7 + es)
    N, n, m, g = 8, 6, 6, 1
    alpha = rbits(rng, N, n)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = _dpf_wide(dev, enc(alpha, es), None, s0, s1, n, m, g, es, N)
    x = rbits(rng, N, 8 * es)
    size = 1 << n
    outall = dev.empty(N * es * size)
    wcall(dev, "dpfEvalAll", outall, N * es * size, dev.put(enc(x, es)), N * es, key, key.numel(), dev.put(s0), N * 16,
          0, n, m, g, es, N, None, 0)
    allv = dec(dget(dev, outall), es)
    points = rbits(rng, 5, 8 * es)
    outm = dev.empty(N * es * len(points))
    wcall(dev, "dpfEvalMulti", outm, N * es * len(points), dev.put(enc(x, es)), N * es, key, key.numel(),
          dev.put(s0), N * 16, 0, dev.put(enc(points, es)), es * len(points), n, m, g, es, N, None, 0)
    multi = dec(dget(dev, outm), es)
    M = (1 << (8 * es)) - 1
    for j in range(size):
        out = dev.empty(N * es)
        wcall(dev, "dpfEval", out, N * es, dev.put(enc([j] * N, es)), N * es, key, key.numel(), dev.put(s0), N * 16, 0,
              n, m, g, es, N, None, 0)
        at_j = dec(dget(dev, out), es)
        for i in range(N):
            assert allv[i * size + (x[i] - j) % size] == at_j[i], (i, j)
    for p, pt in enumerate(points):
        out = dev.empty(N * es)
        xs = [(xi - pt) & M for xi in x]
        wcall(dev, "dpfEval", out, N * es, dev.put(enc(xs, es)), N * es, key, key.numel(), dev.put(s0), N * 16, 0, n,
              m, g, es, N, None, 0)
        at_p = dec(dget(dev, out), es)
        for i in range(N):
            assert multi[i * len(points) + p] == at_p[i], (i, p)


# ---------------------------------------------------------------- Grotto ----

@pytest.mark.parametrize("es", WIDE)
def test_grotto_low64_and_full_width(dev, es):
    rng = np.random.default_rng(200 + es)
    N, n = 200, 40
    alpha = rng.integers(0, 2**n, N, dtype=np.uint64)
    x = rng.integers(0, 2**n, N, dtype=np.uint64)
    x[::5] = alpha[::5]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kn = dev.get(grotto_keygen(dev, alpha, s0, s1, n, 8)).view(np.uint8)
    ks = key_size(dev, "grottoGetKeyDataSize", n, es, N)
    key = dev.empty(ks)
    wcall(dev, "grottoKeyGen", key, ks, dev.put(widen(alpha, es)), N * es, dev.put(s0), N * 16, dev.put(s1), N * 16,
          n, es, N)
    kw = dget(dev, key)
    head = N * 16 * (n - 6) + N * 8
    _tail_low64_equal(kw, kn, head, 2 * N, 2 * N, es)
    for party, seed in ((0, s0), (1, s1)):
        for eb in (False, True):
            yn = dev.get(grotto_eval(dev, x, dev.put(kn), seed, party, n, 8, eb))
            out = dev.empty(N * es)
            wcall(dev, "grottoEval", out, N * es, dev.put(widen(x, es)), N * es, key, ks, dev.put(seed), N * 16, eb,
                  party, n, es, N, None, 0)
            assert np.array_equal(low64(dget(dev, out), es), yn.view(np.uint64)), (party, eb)
        yn = dev.get(grotto_eq_eval(dev, x, dev.put(kn), seed, party, n, 8))
        out = dev.empty(N * es)
        wcall(dev, "grottoEqEval", out, N * es, dev.put(widen(x, es)), N * es, key, ks, dev.put(seed), N * 16, party,
              n, es, N, None, 0)
        assert np.array_equal(low64(dget(dev, out), es), yn.view(np.uint64)), party
    # full width: n = 8 es; XOR of the shares is [x < alpha], [x <= alpha], [x == alpha]
    n, N = 8 * es, 64
    alpha = rbits(rng, N, n)
    x = rbits(rng, N, n)
    x[::4] = alpha[::4]
    x[1::4] = [a >> 1 for a in alpha[1::4]]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ks = key_size(dev, "grottoGetKeyDataSize", n, es, N)
    key = dev.empty(ks)
    wcall(dev, "grottoKeyGen", key, ks, dev.put(enc(alpha, es)), N * es, dev.put(s0), N * 16, dev.put(s1), N * 16, n,
          es, N)
    res = {}
    for name, eb, fn in (("lt", False, "grottoEval"), ("le", True, "grottoEval"), ("eq", None, "grottoEqEval")):
        ys = []
        for party, seed in ((0, s0), (1, s1)):
            out = dev.empty(N * es)
            if eb is None:
                wcall(dev, fn, out, N * es, dev.put(enc(x, es)), N * es, key, ks, dev.put(seed), N * 16, party, n, es,
                      N, None, 0)
            else:
                wcall(dev, fn, out, N * es, dev.put(enc(x, es)), N * es, key, ks, dev.put(seed), N * 16, eb, party, n,
                      es, N, None, 0)
            ys.append(dec(dget(dev, out), es))
        res[name] = [a ^ b for a, b in zip(*ys)]
    for i in range(N):
        assert res["lt"][i] == int(x[i] < alpha[i]), i
        assert res["le"][i] == int(x[i] <= alpha[i]), i
        assert res["eq"][i] == int(x[i] == alpha[i]), i


def _signed(v, bits):
    return v - (1 << bits) if v >> (bits - 1) else v


@pytest.mark.parametrize("es", [16, 32])
def test_grotto_mic_and_lut_full_width(dev, es):
    """Over n = 8 es, Grotto MIC shares XOR to the signed n-bit containment of the unmasked x, and the
    interval LUT reconstructs E * T to LUT[matched interval], on the signed n-bit range cut into I
    contiguous intervals (tests/grotto/grotto_interval_lut_eval_test_case.h)."""
    rng = np.random.default_rng(300 + es)
    n, N, I, lutNum = 8 * es, 48, 3, 2
    M, Mn = (1 << (8 * es)) - 1, (1 << n) - 1
    dom, umin = 1 << n, 1 << (n - 1)
    left = [(umin + dom * k // I) & Mn for k in range(I)]
    right = [(umin + dom * (k + 1) // I - 1) & Mn for k in range(I)]
    alpha = rbits(rng, N, n)
    x = rbits(rng, N, n)
    x[:I], x[I:2 * I] = left, right
    masked = [(a + b) & Mn for a, b in zip(x, alpha)]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ks = key_size(dev, "grottoGetKeyDataSize", n, es, N)
    key = dev.empty(ks)
    wcall(dev, "grottoKeyGen", key, ks, dev.put(enc(alpha, es)), N * es, dev.put(s0), N * 16, dev.put(s1), N * 16, n,
          es, N)
    lut = rbits(rng, I * lutNum, 8 * es)
    mic, E, T = [], [], []
    for party, seed in ((0, s0), (1, s1)):
        out = dev.empty(N * es * I)
        wcall(dev, "grottoMICEval", out, N * es * I, dev.put(enc(masked, es)), N * es, key, ks, dev.put(seed), N * 16,
              party, dev.put(enc(left, es)), I * es, dev.put(enc(right, es)), I * es, n, es, N, None, 0)
        mic.append(dec(dget(dev, out), es))
        oe, ot = dev.empty(N * es), dev.empty(N * es * lutNum)
        wcall(dev, "grottoIntervalLutEval", oe, N * es, ot, N * es * lutNum, dev.put(enc(masked, es)), N * es, key, ks,
              dev.put(seed), N * 16, party, dev.put(enc(left, es)), I * es, dev.put(enc(right, es)), I * es,
              dev.put(enc(lut, es)), I * lutNum * es, n, n, es, N, None, 0)
        E.append(dec(dget(dev, oe), es))
        T.append(dec(dget(dev, ot), es))
    for i in range(N):
        xs = _signed(x[i], n)
        hits = [k for k in range(I) if _signed(left[k], n) <= xs <= _signed(right[k], n)]
        assert len(hits) == 1, i
        for k in range(I):
            assert (mic[0][i * I + k] ^ mic[1][i * I + k]) & 1 == int(k in hits), (i, k)
        e = (E[0][i] + E[1][i]) & M
        assert e in (1, M), i
        for j in range(lutNum):
            t = (T[0][i * lutNum + j] + T[1][i * lutNum + j]) & M
            assert (e * t) & M == lut[hits[0] + j * I], (i, j)


# ------------------------------------------------------------------- MIC ----

@pytest.mark.parametrize("es", WIDE)
def test_mic_low64_and_full_width(dev, es):
    rng = np.random.default_rng(400 + es)
    N, n, m, I = 150, 40, 40, 3
    M = (1 << (8 * es)) - 1
    cuts = sorted(int(v) for v in rng.integers(1, 2**(n - 2), I * 2, dtype=np.uint64))
    left = np.array(cuts[0::2], dtype=np.uint64)
    right = np.array(cuts[1::2], dtype=np.uint64)
    alpha = rng.integers(0, 2**n, N, dtype=np.uint64)
    x = rng.integers(0, 2**n, N, dtype=np.uint64)
    masked = (x + alpha) & np.uint64(2**n - 1)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kn, zn = mic_keygen(dev, alpha, s0, s1, left, right, n, m, 8)
    kn, zn = dev.get(kn).view(np.uint8), dev.get(zn)
    ks = key_size(dev, "dcfMICGetKeyDataSize", n, m, es, N)
    key, z = dev.empty(ks), dev.empty(N * I * es)
    wcall(dev, "dcfMICKeyGen", key, ks, z, N * I * es, dev.put(widen(alpha, es)), N * es, dev.put(s0), N * 16,
          dev.put(s1), N * 16, dev.put(widen(left, es)), I * es, dev.put(widen(right, es)), I * es, n, m, es, N)
    kw, zw = dget(dev, key), dget(dev, z)
    _tail_low64_equal(kw, kn, N * 16 * n, N * (n + 3), N * (n + 3), es)
    assert np.array_equal(low64(zw, es), zn.view(np.uint64))
    z0 = rng.integers(0, 2**63, N * I, dtype=np.uint64)
    zw_vals = dec(zw, es)
    z0w = [int(v) for v in z0]
    z1w = [(a - b) & M for a, b in zip(zw_vals, z0w)]
    z1n = ((zn.view(np.uint64) - z0) & np.uint64(2**64 - 1))
    for party, seed, zz, zzw in ((0, s0, z0, z0w), (1, s1, z1n, z1w)):
        yn = dev.get(mic_eval(dev, masked, dev.put(kn), zz, seed, party, left, right, n, m, 8))
        out = dev.empty(N * I * es)
        wcall(dev, "dcfMICEval", out, N * I * es, dev.put(widen(masked, es)), N * es, key, ks, dev.put(enc(zzw, es)),
              N * I * es, dev.put(seed), N * 16, party, dev.put(widen(left, es)), I * es, dev.put(widen(right, es)),
              I * es, n, m, es, N, None, 0)
        assert np.array_equal(low64(dget(dev, out), es), yn.view(np.uint64)), party
    # full width: [x in [L, R]] over n = m = 8 es
    n = m = 8 * es
    N = 48
    cuts = sorted(rbits(rng, 2 * I, n - 1))
    left, right = cuts[0::2], cuts[1::2]
    alpha = rbits(rng, N, n)
    x = rbits(rng, N, n - 1)
    x[:I], x[I:2 * I] = left, right
    masked = [(a + b) % (1 << n) for a, b in zip(x, alpha)]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ks = key_size(dev, "dcfMICGetKeyDataSize", n, m, es, N)
    key, z = dev.empty(ks), dev.empty(N * I * es)
    wcall(dev, "dcfMICKeyGen", key, ks, z, N * I * es, dev.put(enc(alpha, es)), N * es, dev.put(s0), N * 16,
          dev.put(s1), N * 16, dev.put(enc(left, es)), I * es, dev.put(enc(right, es)), I * es, n, m, es, N)
    zv = dec(dget(dev, z), es)
    z0 = rbits(rng, N * I, 8 * es)
    z1 = [(a - b) & M for a, b in zip(zv, z0)]
    ys = []
    for party, seed, zz in ((0, s0, z0), (1, s1, z1)):
        out = dev.empty(N * I * es)
        wcall(dev, "dcfMICEval", out, N * I * es, dev.put(enc(masked, es)), N * es, key, ks, dev.put(enc(zz, es)),
              N * I * es, dev.put(seed), N * 16, party, dev.put(enc(left, es)), I * es, dev.put(enc(right, es)),
              I * es, n, m, es, N, None, 0)
        ys.append(dec(dget(dev, out), es))
    for i in range(N):
        for q in range(I):
            assert (ys[0][i * I + q] + ys[1][i * I + q]) & M == int(left[q] <= x[i] <= right[q]), (i, q)


# ---------------------------------------------------------------- spline ----

@pytest.mark.parametrize("es", [16, 24, 32])
def test_spline_low64_and_full_width(dev, es):
    rng = np.random.default_rng(500 + es)
    d, I = 2, 3
    C, g = d + 1, I * (d + 1)
    M = (1 << (8 * es)) - 1
    N, n, m = 120, 48, 48
    cut = sorted(int(v) for v in rng.integers(1, 2**n - 2, I - 1, dtype=np.uint64))
    left = np.array([0] + [c + 1 for c in cut], dtype=np.uint64)
    right = np.array(cut + [2**n - 1], dtype=np.uint64)
    coeff = rng.integers(0, 2**m, g, dtype=np.uint64)
    alpha = rng.integers(0, 2**n, N, dtype=np.uint64)
    x = rng.integers(0, 2**n, N, dtype=np.uint64)
    masked = (x + alpha) & np.uint64(2**n - 1)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kn, en, bn = (dev.get(t) for t in spline_keygen(dev, alpha, s0, s1, coeff, d, left, right, n, m, 8))
    ks = key_size(dev, "dcfSplineGetKeyDataSize", d, I, n, m, es, N)
    key, e, b = dev.empty(ks), dev.empty(N * g * es), dev.empty(N * g * es)
    wcall(dev, "dcfSplineKeyGen", key, ks, e, N * g * es, b, N * g * es, dev.put(widen(alpha, es)), N * es,
          dev.put(s0), N * 16, dev.put(s1), N * 16, dev.put(widen(coeff, es)), g * es, d, dev.put(widen(left, es)),
          I * es, dev.put(widen(right, es)), I * es, I, n, m, es, N)
    kw = dget(dev, key)
    _tail_low64_equal(kw, kn.view(np.uint8), N * 16 * n, N * (n * g + 2 + g), N * (n * g + 2 + g), es)
    assert np.array_equal(low64(dget(dev, e), es), en.view(np.uint64))
    assert np.array_equal(low64(dget(dev, b), es), bn.view(np.uint64))
    cs = key_size(dev, "dcfSplineGetCacheDataSize", d, I, n, m, es, N)
    for party, seed in ((0, s0), (1, s1)):
        yn = dev.get(spline_eval(dev, masked, dev.put(kn), en, bn, seed, party, left, right, d, n, m, 8))
        out, cache = dev.empty(N * es), dev.empty(cs)
        wcall(dev, "dcfSplineEval", out, N * es, dev.put(widen(masked, es)), N * es, key, ks, e, N * g * es, b,
              N * g * es, dev.put(seed), N * 16, party, dev.put(widen(left, es)), I * es, dev.put(widen(right, es)),
              I * es, I, d, n, m, es, N, cache, cs)
        assert np.array_equal(low64(dget(dev, out), es), yn.view(np.uint64)), party
    # full width, n = m = 8 es: shares reconstruct to f_q(x) mod 2^m, coefficient shares e/beta as dealt
    n = m = 8 * es
    N = 32
    cut = sorted(rbits(rng, I - 1, n))
    left = [0] + [c + 1 for c in cut]
    right = cut + [(1 << n) - 1]
    coeff = rbits(rng, g, m)
    alpha = rbits(rng, N, n)
    x = rbits(rng, N, n)
    masked = [(a + b) % (1 << n) for a, b in zip(x, alpha)]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ks = key_size(dev, "dcfSplineGetKeyDataSize", d, I, n, m, es, N)
    key, e, b = dev.empty(ks), dev.empty(N * g * es), dev.empty(N * g * es)
    wcall(dev, "dcfSplineKeyGen", key, ks, e, N * g * es, b, N * g * es, dev.put(enc(alpha, es)), N * es, dev.put(s0),
          N * 16, dev.put(s1), N * 16, dev.put(enc(coeff, es)), g * es, d, dev.put(enc(left, es)), I * es,
          dev.put(enc(right, es)), I * es, I, n, m, es, N)
    ev, bv = dec(dget(dev, e), es), dec(dget(dev, b), es)
    e0, b0 = rbits(rng, N * g, 8 * es), rbits(rng, N * g, 8 * es)
    shares = ((e0, b0), ([(p - q) & M for p, q in zip(ev, e0)], [(p - q) & M for p, q in zip(bv, b0)]))
    cs = key_size(dev, "dcfSplineGetCacheDataSize", d, I, n, m, es, N)
    ys = []
    for party, seed in ((0, s0), (1, s1)):
        out, cache = dev.empty(N * es), dev.empty(cs)
        es_, bs_ = shares[party]
        wcall(dev, "dcfSplineEval", out, N * es, dev.put(enc(masked, es)), N * es, key, ks, dev.put(enc(es_, es)),
              N * g * es, dev.put(enc(bs_, es)), N * g * es, dev.put(seed), N * 16, party, dev.put(enc(left, es)),
              I * es, dev.put(enc(right, es)), I * es, I, d, n, m, es, N, cache, cs)
        ys.append(dec(dget(dev, out), es))
    for i in range(N):
        q = max(k for k in range(I) if left[k] <= x[i])
        want = sum(coeff[q * C + k] * x[i] ** k for k in range(C)) % (1 << m)
        assert (ys[0][i] + ys[1][i]) % (1 << m) == want, i


# ------------------------------------------------------------------ OTTT ----

@pytest.mark.parametrize("es", WIDE)
def test_ottt_low64_and_full_width(dev, es):
    rng = np.random.default_rng(600 + es)
    N, n, lutNum = 40, 8, 2
    M = (1 << (8 * es)) - 1
    size = 1 << n
    alpha = rng.integers(0, size, N, dtype=np.uint64)
    mask_key = rng.integers(0, 256, N * size // 8, dtype=np.uint8)
    x = rng.integers(0, size, N, dtype=np.uint64)
    masked = (x + alpha) & np.uint64(size - 1)
    lut = rbits(rng, lutNum * size, 8 * es)
    keys = []
    for ees, a in ((8, alpha), (es, widen(alpha, es))):
        kk = dev.put(mask_key.copy())
        wcall(dev, "otttKeyGen", kk, kk.numel(), dev.put(a), N * ees, n, ees, N)
        keys.append(dget(dev, kk))
    assert np.array_equal(keys[0], keys[1])
    k0 = mask_key  # party 0 holds the random mask, party 1 the mask with bit alpha flipped
    E, T = [], []
    for party, kb in ((0, k0), (1, keys[1])):
        oe, ot = dev.empty(N * es), dev.empty(N * es * lutNum)
        wcall(dev, "otttLutEval", oe, N * es, ot, N * es * lutNum, dev.put(widen(masked, es)), N * es, dev.put(kb),
              kb.size, party, dev.put(enc(lut, es)), lutNum * size * es, n, es, N)
        E.append(dec(dget(dev, oe), es))
        T.append(dec(dget(dev, ot), es))
        lut64 = np.array([v & (2**64 - 1) for v in lut], dtype=np.uint64)
        en, tn = ottt_eval(dev, masked, dev.put(kb), party, lut64, n, 8)
        assert [v & (2**64 - 1) for v in E[-1]] == [int(v) for v in dev.get(en).view(np.uint64)], party
        assert [v & (2**64 - 1) for v in T[-1]] == [int(v) for v in dev.get(tn).view(np.uint64)], party
    for i in range(N):
        e = (E[0][i] + E[1][i]) & M
        for j in range(lutNum):
            t = (T[0][i * lutNum + j] + T[1][i * lutNum + j]) & M
            assert (e * t) & M == lut[int(x[i]) + j * size], (i, j)
