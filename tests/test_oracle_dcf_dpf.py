"""Pins the oracle's DCF and DPF against the mathematics and independent AES.

The reference ships no golden vectors and its CPU path cannot be built here
(its wideint submodule is empty), so each oracle function is pinned by checks
that a plausible mistake would fail:

* reconstruction: share0 + share1 == beta*[x < alpha] (DCF) / beta*[x == alpha]
  (DPF), brute force over every (alpha, x) for tiny widths and random inputs on
  the reference's own test grid (tests/dcf/dcf_eval_cpu.cpp:49-193,
  tests/dpf/dpf_eval_case.h:70-92);
* first-level correction words recomputed with the system libcrypto AES: the
  PRG key (seed with bit 0 of byte 0 cleared), the counter plaintexts, the
  t-bit split, the keep/lose choice from alpha's top bit, tLCW/tRCW bit 0 and
  vCW[0] including the value-conversion byte rules (direct bytes, AES refill
  keyed by v with (0, element-index) counters, src/impl/dcf.h:69-137) and the
  DPF conversion argument swap (src/impl/dpf.h:74,278,283);
* the key byte layout: element i's correction words sit at the offsets of
  src/impl/dcf.h:139-167 / src/impl/dpf.h:130-156 (checked on i > 0);
* a depth-1 evaluation recomputed from libcrypto blocks, which pins the eval
  side's plaintexts, sign and unreduced output (src/impl/dcf.h:467-490);
* error codes of src/kernel/dcf.h:115-163 and src/kernel/dpf.h:115-154.
"""
import itertools

import numpy as np
import pytest

from fss_testlib import (DT, Aes, OracleBackend, clear_lsb, dcf_eval, dcf_keygen, dpf_eval, dpf_eval_all,
                         dpf_eval_multi, dpf_keygen, le_block, mask, ocall, rand_elems, rand_seeds, xor16)

BE = OracleBackend()
need_aes = pytest.mark.skipif(not Aes.available(), reason="system libcrypto not found")


def _rec(y0, y1, m, g):
    return ((y0.astype(np.uint64) + y1.astype(np.uint64)) & np.uint64(mask(m))).reshape(-1, g)


# ---------------------------------------------------------------- DCF -----


@pytest.mark.parametrize("n,es", [(1, 1), (3, 1), (5, 2), (6, 8)])
@pytest.mark.parametrize("g", [1, 3])
def test_dcf_bruteforce(n, es, g):
    rng = np.random.default_rng(n * 10 + g)
    m = 8 * es
    pairs = np.array(list(itertools.product(range(1 << n), range(1 << n))), dtype=np.uint64)
    alpha = pairs[:, 0].astype(DT[es])
    x = pairs[:, 1].astype(DT[es])
    N = len(alpha)
    beta = rand_elems(rng, N * g, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dcf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
    y0 = dcf_eval(BE, x, key, s0, 0, n, m, g, es)
    y1 = dcf_eval(BE, x, key, s1, 1, n, m, g, es)
    want = np.where((x < alpha)[:, None], beta.reshape(-1, g).astype(np.uint64), 0)
    assert np.array_equal(_rec(y0, y1, m, g), want)


GRID = {1: [1, 7, 8], 2: [1, 15, 16], 4: [1, 31, 32], 8: [1, 63, 64]}


@pytest.mark.parametrize("es", [1, 2, 4, 8])
def test_dcf_reference_grid(es):
    """The reference test grid: (n, m) in {1, w-1, w}^2, groupSize 1..4."""
    rng = np.random.default_rng(es)
    for n, m in itertools.product(GRID[es], GRID[es]):
        for g, N in [(1, 1023), (2, 8), (3, 1), (4, 64)]:
            alpha = rand_elems(rng, N, n, es)
            x = rand_elems(rng, N, n, es)
            beta = rand_elems(rng, N * g, m, es)
            masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
            s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
            key = dcf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
            y0 = dcf_eval(BE, masked, key, s0, 0, n, m, g, es)
            y1 = dcf_eval(BE, masked, key, s1, 1, n, m, g, es)
            want = np.where((masked < alpha)[:, None], beta.reshape(-1, g).astype(np.uint64), 0)
            assert np.array_equal(_rec(y0, y1, m, g), want), (n, m, g)


def test_dcf_cpu_workload_n16_2pow16():
    """The reference's CPU-sized DCF workload: bitWidthIn = bitWidthOut = 16,
    2^16 elements stored in 64-bit elements, beta = NULL (= 1), alpha and the
    masked input uniform over [0, 2^16); reconstruction (s0 + s1) mod 2^16 ==
    [x < alpha]."""
    rng = np.random.default_rng(16)
    N, n, m, es = 1 << 16, 16, 16, 8
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dcf_keygen(BE, alpha, None, s0, s1, n, m, 1, es)
    assert key.nbytes == N * (16 * n + es * n + 3 * es)
    y0 = dcf_eval(BE, x, key, s0, 0, n, m, 1, es)
    y1 = dcf_eval(BE, x, key, s1, 1, n, m, 1, es)
    assert np.array_equal(_rec(y0, y1, m, 1).reshape(-1), (x < alpha).astype(np.uint64))


def test_dcf_beta_null_is_one():
    rng = np.random.default_rng(5)
    N, n, es = 64, 8, 1
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dcf_keygen(BE, alpha, None, s0, s1, n, 8, 1, es)
    y0 = dcf_eval(BE, x, key, s0, 0, n, 8, 1, es)
    y1 = dcf_eval(BE, x, key, s1, 1, n, 8, 1, es)
    assert np.array_equal((y0 + y1).astype(np.uint8), (x < alpha).astype(np.uint8))


def _children(seed: bytes, count: int):
    """AES_{seed & ~1}(u64 2k || u64 2k+1), k < count (split t-bit later)."""
    k = clear_lsb(seed)
    return [Aes.encrypt(k, le_block(2 * j, 2 * j + 1)) for j in range(count)]


def _conv(v: bytes, g: int, bits: int):
    """Value conversion of the reference's DcfConvertCtx, from its definition."""
    nb = (bits + 7) // 8
    if nb * g <= 16:
        return [int.from_bytes(v[j * nb:(j + 1) * nb], "little") for j in range(g)]
    out, buf = [], None
    for j in range(g):
        val = 0
        for i in range(nb):
            idx = (j * nb + i) % 16
            if idx == 0:
                buf = Aes.encrypt(v, le_block(0, j))
            val |= buf[idx] << (8 * i)
        out.append(val)
    return out


def _dcf_tail(key, n, g, es, i, N):
    base = N * 16 * n + i * (es * n * g + 2 * es + es * g)
    raw = key[base:base + es * (n * g + 2 + g)].view(DT[es])
    return raw[:n * g], int(raw[n * g]), int(raw[n * g + 1]), raw[n * g + 2:]


@need_aes
@pytest.mark.parametrize("n,m,g,es", [(8, 8, 1, 1), (32, 32, 1, 4), (16, 16, 2, 2), (32, 32, 5, 4), (24, 24, 6, 4),
                                      (64, 64, 3, 8), (12, 20, 1, 4)])
def test_dcf_first_level_closed_form(n, m, g, es):
    rng = np.random.default_rng(n + 100 * g)
    N = 3
    alpha = rand_elems(rng, N, n, es)
    beta = rand_elems(rng, N * g, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dcf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
    W = mask(8 * es)
    for i in range(N):
        c0 = _children(s0[16 * i:16 * i + 16].tobytes(), 4)
        c1 = _children(s1[16 * i:16 * i + 16].tobytes(), 4)
        keep = (int(alpha[i]) >> (n - 1)) & 1
        lose = 1 - keep
        sl0, sl1 = clear_lsb(c0[2 * lose]), clear_lsb(c1[2 * lose])
        assert key[16 * n * i:16 * n * i + 16].tobytes() == xor16(sl0, sl1)
        vcw, tl, tr, _ = _dcf_tail(key, n, g, es, i, N)
        tL0, tR0, tL1, tR1 = c0[0][0] & 1, c0[2][0] & 1, c1[0][0] & 1, c1[2][0] & 1
        assert tl & 1 == tL0 ^ tL1 ^ keep ^ 1
        assert tr & 1 == tR0 ^ tR1 ^ keep
        # party 1 starts with t = 1, so vCW[0] = -(conv(vLose1) - conv(vLose0)) - [lose is L]*beta
        a = _conv(c1[2 * lose + 1], g, m)
        b = _conv(c0[2 * lose + 1], g, m)
        for j in range(g):
            want = -(a[j] - b[j])
            if lose == 0:
                want -= int(beta[i * g + j])
            assert int(vcw[j]) == want & W, (i, j)


@need_aes
@pytest.mark.parametrize("es,m", [(1, 8), (4, 32), (4, 24), (8, 64)])
def test_dcf_depth1_eval_closed_form(es, m):
    """n = 1: out_b = sgn_b*(conv(v_x) + t_b*vCW) + sgn_b*(conv(s') + t'*lastCW), unreduced."""
    rng = np.random.default_rng(es + m)
    N, n, g = 8, 1, 1
    alpha = np.array([0, 1, 0, 1, 1, 0, 1, 0], dtype=DT[es])
    x = np.array([0, 0, 1, 1, 0, 1, 1, 0], dtype=DT[es])
    beta = rand_elems(rng, N, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dcf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
    W = mask(8 * es)
    for party, seeds in ((0, s0), (1, s1)):
        y = dcf_eval(BE, x, key, seeds, party, n, m, g, es)
        for i in range(N):
            c = _children(seeds[16 * i:16 * i + 16].tobytes(), 4)
            vcw, tl, tr, last = _dcf_tail(key, n, g, es, i, N)
            sCW = key[16 * i:16 * i + 16].tobytes()
            xb = int(x[i]) & 1
            s_next = clear_lsb(c[2 * xb])
            t_next = c[2 * xb][0] & 1
            if party:  # t = 1 at the root: apply the correction words
                s_next = xor16(s_next, sCW)
                t_next ^= (tr if xb else tl) & 1
            sgn = -1 if party else 1
            acc = sgn * (_conv(c[2 * xb + 1], 1, m)[0] + (int(vcw[0]) if party else 0))
            acc += sgn * (_conv(s_next, 1, m)[0] + (int(last[0]) if t_next else 0))
            assert int(y[i]) == acc & W


def test_dcf_errors():
    z = np.zeros(4096, dtype=np.uint8)
    ks = 1 * (16 * 8 + 1 * 8 + 3)
    # element size first, then bitwidths, party, group, key, out, x, seed, cache
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 3, 1, None, 0) == -11
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 16, 0, 9, 8, 1, 1, 1, None, 0) == -10
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 16, 0, 8, 0, 1, 1, 1, None, 0) == -10
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 16, 2, 8, 8, 1, 1, 1, None, 0) == -12
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 0, 1, 1, None, 0) == -16
    assert ocall("dcfEval", z, 1, z, 1, z, ks + 1, z, 16, 0, 8, 8, 1, 1, 1, None, 0) == -2
    assert ocall("dcfEval", z, 2, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0) == -8
    assert ocall("dcfEval", z, 1, z, 2, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0) == -9
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 15, 0, 8, 8, 1, 1, 1, None, 0) == -5
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, z, 7) == -13
    assert ocall("dcfEval", z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, z, 16 * 8 + 8) == 0
    assert ocall("dcfKeyGen", z, ks, z, 1, None, 0, z, 16, z, 16, 8, 8, 1, 1, 1) == 0
    assert ocall("dcfKeyGen", z, ks, z, 1, z, 2, z, 16, z, 16, 8, 8, 1, 1, 1) == -4
    assert ocall("dcfKeyGen", z, ks, z, 2, None, 0, z, 16, z, 16, 8, 8, 1, 1, 1) == -3
    assert ocall("dcfKeyGen", z, ks, z, 1, None, 0, z, 16, z, 17, 8, 8, 1, 1, 1) == -5


# ---------------------------------------------------------------- DPF -----


@pytest.mark.parametrize("n,es,m", [(1, 1, 8), (4, 1, 8), (6, 2, 16), (5, 4, 32), (4, 8, 64), (3, 4, 17)])
@pytest.mark.parametrize("g", [1, 2])
def test_dpf_bruteforce(n, es, m, g):
    rng = np.random.default_rng(n + es + g)
    pairs = np.array(list(itertools.product(range(1 << n), range(1 << n))), dtype=np.uint64)
    alpha = pairs[:, 0].astype(DT[es])
    x = pairs[:, 1].astype(DT[es])
    N = len(alpha)
    beta = rand_elems(rng, N * g, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dpf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
    y0 = dpf_eval(BE, x, key, s0, 0, n, m, g, es)
    y1 = dpf_eval(BE, x, key, s1, 1, n, m, g, es)
    want = np.where((x == alpha)[:, None], beta.reshape(-1, g).astype(np.uint64), 0)
    assert np.array_equal(_rec(y0, y1, m, g), want)


@pytest.mark.parametrize("es", [1, 2, 4, 8])
def test_dpf_reference_grid(es):
    rng = np.random.default_rng(30 + es)
    for n, m in itertools.product(GRID[es], GRID[es]):
        g, N = (1, 300) if n > 1 else (2, 16)
        alpha = rand_elems(rng, N, n, es)
        x = rand_elems(rng, N, n, es)
        x[::3] = 0
        beta = rand_elems(rng, N * g, m, es)
        masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
        s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
        key = dpf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
        y0 = dpf_eval(BE, masked, key, s0, 0, n, m, g, es)
        y1 = dpf_eval(BE, masked, key, s1, 1, n, m, g, es)
        want = np.where((x == 0)[:, None], beta.reshape(-1, g).astype(np.uint64), 0)
        assert np.array_equal(_rec(y0, y1, m, g), want), (n, m)


@need_aes
@pytest.mark.parametrize("m,es", [(8, 1), (16, 2), (17, 4), (32, 4), (64, 8)])
def test_dpf_depth1_closed_form(m, es):
    """n = 1 pins sCW, the t-bits and lastCW including the conversion argument swap:
    with groupSize 1 the conversion reads ceil(1/8) = 1 byte, directly from the
    seed while bitWidthOut <= 16 and from AES_seed(0,0) beyond."""
    rng = np.random.default_rng(m)
    N, n, g = 6, 1, 1
    alpha = np.array([0, 1, 1, 0, 1, 0], dtype=DT[es])
    beta = rand_elems(rng, N, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dpf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
    W = mask(8 * es)

    def conv(s):
        return s[0] if m <= 16 else Aes.encrypt(s, le_block(0, 0))[0]

    for i in range(N):
        c0 = _children(s0[16 * i:16 * i + 16].tobytes(), 2)
        c1 = _children(s1[16 * i:16 * i + 16].tobytes(), 2)
        keep = int(alpha[i]) & 1
        lose = 1 - keep
        sCW = xor16(clear_lsb(c0[lose]), clear_lsb(c1[lose]))
        assert key[16 * i:16 * i + 16].tobytes() == sCW
        tail = key[N * 16 + i * 3 * es:N * 16 + (i + 1) * 3 * es].view(DT[es])
        tL = (c0[0][0] ^ c1[0][0]) & 1
        tR = (c0[1][0] ^ c1[1][0]) & 1
        tLCW, tRCW = tL ^ keep ^ 1, tR ^ keep
        assert int(tail[0]) == tLCW and int(tail[1]) == tRCW
        # party 0 starts at t=0, party 1 at t=1 (applies the correction)
        sk0 = clear_lsb(c0[keep])
        sk1 = xor16(clear_lsb(c1[keep]), sCW)
        t1 = (c1[keep][0] & 1) ^ (tRCW if keep else tLCW)
        want = (beta[i].item() - conv(sk0) + conv(sk1)) * (-1 if t1 else 1)
        assert int(tail[2]) == want & W


@pytest.mark.parametrize("n,es,g", [(3, 1, 1), (5, 2, 2), (4, 8, 1)])
def test_dpf_eval_all_one_hot(n, es, g):
    """tests/dpf/dpf_eval_all_case.h:80-87: the point x (= maskedX - alpha) holds beta."""
    rng = np.random.default_rng(n * es)
    N, m = 9, 8 * es
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    beta = rand_elems(rng, N * g, m, es)
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dpf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
    y0 = dpf_eval_all(BE, masked, key, s0, 0, n, m, g, es)
    y1 = dpf_eval_all(BE, masked, key, s1, 1, n, m, g, es)
    rec = _rec(y0, y1, m, g).reshape(N, 1 << n, g)
    want = np.zeros_like(rec)
    for i in range(N):
        want[i, int(x[i])] = beta.reshape(-1, g)[i]
    assert np.array_equal(rec, want)
    # and each slot equals the single-point evaluation at j
    for i in range(2):
        for j in range(1 << n):
            yj = dpf_eval(BE, np.array([j], dtype=DT[es]), _dpf_key_slice(key, n, g, es, i, N), s0[16 * i:16 * i + 16],
                          0, n, m, g, es)
            idx = (int(masked[i]) - j) % (1 << n)
            assert np.array_equal(y0.reshape(N, 1 << n, g)[i, idx], yj)


def _dpf_key_slice(key, n, g, es, i, N):
    head = key[16 * n * i:16 * n * (i + 1)]
    base = N * 16 * n + i * es * (2 + g)
    return np.concatenate([head, key[base:base + es * (2 + g)]])


def test_dpf_eval_multi_matches_points():
    rng = np.random.default_rng(77)
    N, n, m, g, es = 5, 12, 16, 2, 2
    alpha = rand_elems(rng, N, n, es)
    beta = rand_elems(rng, N * g, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = dpf_keygen(BE, alpha, beta, s0, s1, n, m, g, es)
    x = rand_elems(rng, N, n, es)
    pts = np.concatenate([rand_elems(rng, 6, n, es), np.array([0], dtype=DT[es])])
    y0 = dpf_eval_multi(BE, x, key, s0, 0, pts, n, m, g, es).reshape(N, len(pts), g)
    y1 = dpf_eval_multi(BE, x, key, s1, 1, pts, n, m, g, es).reshape(N, len(pts), g)
    rec = (y0.astype(np.uint64) + y1) & np.uint64(mask(m))
    for i in range(N):
        for j, p in enumerate(pts):
            hit = ((int(x[i]) - int(p)) & mask(n)) == int(alpha[i])
            want = beta.reshape(-1, g)[i].astype(np.uint64) if hit else np.zeros(g, np.uint64)
            assert np.array_equal(rec[i, j], want)


def test_dpf_errors():
    z = np.zeros(4096, dtype=np.uint8)
    ks = 16 * 8 + 3
    assert ocall("dpfKeyGen", z, ks, z, 1, None, 0, z, 16, z, 16, 8, 8, 2, 1, 1) == -2
    ks2 = 16 * 8 + 4
    assert ocall("dpfKeyGen", z, ks2, z, 1, None, 0, z, 16, z, 16, 8, 8, 2, 1, 1) == -4  # beta needed for g > 1
    assert ocall("dpfEval", z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0) == 0
    assert ocall("dpfEval", z, 1, z, 1, z, ks, z, 15, 0, 8, 8, 1, 1, 1, None, 0) == -5
    assert ocall("dpfEval", z, 1, z, 2, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0) == -9
    assert ocall("dpfEval", z, 2, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, None, 0) == -8
    assert ocall("dpfEval", z, 1, z, 1, z, ks, z, 16, 0, 8, 8, 1, 1, 1, z, 5) == -13
    assert ocall("dpfEvalMulti", z, 1, z, 1, z, ks, z, 16, 0, z, 3, 8, 8, 1, 2, 1, None, 0) == -14
