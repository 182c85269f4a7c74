"""Pins the oracle's Grotto, DCF-MIC, DCF-spline and OTTT.

Checks (none of which re-derives a share with the oracle's own code):

* Grotto: XOR reconstruction == [x < alpha] / [x <= alpha] / [x == alpha] by
  brute force (tests/grotto/grotto_eval_test_case.h:62-66), the leaf word
  lastCW == s0.hi ^ s1.hi ^ (1 << (alpha & 63)) at depth 0 (n = 6) and the first
  level's sCW from libcrypto (src/impl/grotto.h:111-205); MIC and interval LUT
  against their plaintext definitions (tests/grotto/grotto_mic_eval_test_case.h).
* DCF-MIC: share0 + share1 == [x in [p, q]] for signed n-bit intervals
  (tests/mic/mic_eval_test_case.h:84-106) and z from its closed form.
* Grotto interval LUT: E*T == LUT[matched interval] (0 where none matches) on
  the reference's signed full partition and on partitions with gaps, with E a
  sign (+1/-1) and T == E*LUT (tests/grotto/grotto_interval_lut_eval_cpu.cpp:
  107-116, grotto_interval_lut_eval_test_case.h:71-119).
* Spline: reconstruction == f_i(x) for the interval holding x
  (tests/spline/spline_eval_test_case.h:128-152); the shifted coefficients beta
  satisfy f(y - alpha) == sum beta_k y^k at random y; e == cr * beta. For
  bitWidthIn != bitWidthOut the reconstruction is f_q(X - alpha) mod 2^out at
  the integer difference X - alpha, X = maskedX mod 2^in, q the interval holding
  (X - alpha) mod 2^in (src/impl/spline.h:96,159-161,206,252).
* OTTT: E*T == LUT[x] (tests/ottt/ottt_eval_cpu.cpp:83-92) and the key differs
  from the caller's mask only in bit alpha.
"""
import itertools

import numpy as np
import pytest

from fss_testlib import (DT, Aes, OracleBackend, clear_lsb, grotto_eq_eval, grotto_eval, grotto_keygen,
                         grotto_lut_eval, grotto_mic_eval, le_block, mask, mic_eval, mic_keygen, ocall, ottt_eval,
                         ottt_keygen, rand_elems, rand_seeds, signed_nbit, spline_eval, spline_keygen, xor16)

BE = OracleBackend()
need_aes = pytest.mark.skipif(not Aes.available(), reason="system libcrypto not found")


# ------------------------------------------------------------- Grotto -----


@pytest.mark.parametrize("n,es", [(6, 1), (7, 1), (8, 1), (9, 2), (10, 8)])
def test_grotto_bruteforce(n, es):
    rng = np.random.default_rng(n)
    pairs = np.array(list(itertools.product(range(1 << n), range(1 << n))), dtype=np.uint64)
    if len(pairs) > 40000:
        pairs = pairs[rng.choice(len(pairs), 40000, replace=False)]
    alpha = pairs[:, 0].astype(DT[es])
    x = pairs[:, 1].astype(DT[es])
    N = len(alpha)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = grotto_keygen(BE, alpha, s0, s1, n, es)
    for eb, want in ((False, x < alpha), (True, x <= alpha)):
        y0 = grotto_eval(BE, x, key, s0, 0, n, es, eb)
        y1 = grotto_eval(BE, x, key, s1, 1, n, es, eb)
        assert np.array_equal((y0 ^ y1) & 1, want.astype(DT[es])), eb
        assert set(np.unique(y0)).issubset({0, 1})
    e0 = grotto_eq_eval(BE, x, key, s0, 0, n, es)
    e1 = grotto_eq_eval(BE, x, key, s1, 1, n, es)
    assert np.array_equal(e0 ^ e1, (x == alpha).astype(DT[es]))


@need_aes
def test_grotto_leaf_and_first_level():
    rng = np.random.default_rng(11)
    N, es = 4, 2
    for n in (6, 9):
        alpha = rand_elems(rng, N, n, es)
        s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
        key = grotto_keygen(BE, alpha, s0, s1, n, es)
        L = n - 6
        for i in range(N):
            a0 = clear_lsb(s0[16 * i:16 * i + 16].tobytes())
            a1 = clear_lsb(s1[16 * i:16 * i + 16].tobytes())
            leaf = key[N * 16 * L + 8 * i:N * 16 * L + 8 * i + 8].tobytes()
            if n == 6:
                want = int.from_bytes(a0[8:], "little") ^ int.from_bytes(a1[8:], "little") ^ (1 << (int(alpha[i]) & 63))
                assert int.from_bytes(leaf, "little") == want
            else:
                lose = 1 - ((int(alpha[i]) >> (n - 1)) & 1)
                c0 = Aes.encrypt(a0, le_block(2 * lose, 2 * lose + 1))
                c1 = Aes.encrypt(a1, le_block(2 * lose, 2 * lose + 1))
                assert key[16 * L * i:16 * L * i + 16].tobytes() == xor16(clear_lsb(c0), clear_lsb(c1))


def _intervals(rng, I, n, es):
    """Signed n-bit intervals with left <= right (tests/mic/mic_eval_test_case.h:70-78)."""
    lo = rand_elems(rng, I, n, es)
    hi = rand_elems(rng, I, n, es)
    L, R = [], []
    for a, b in zip(lo, hi):
        a, b = int(a), int(b)
        if signed_nbit(b, n) < signed_nbit(a, n):
            a, b = b, a
        L.append(a)
        R.append(b)
    return np.array(L, dtype=DT[es]), np.array(R, dtype=DT[es])


@pytest.mark.parametrize("n,es", [(8, 1), (12, 2), (20, 4)])
def test_grotto_mic_and_lut(n, es):
    rng = np.random.default_rng(n + 5)
    N, I = 200, 5
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    left, right = _intervals(rng, I, n, es)
    left[2] = (int(right[1]) + 1) & mask(n)  # exercise the adjacent-interval reuse
    if signed_nbit(int(left[2]), n) > signed_nbit(int(right[2]), n):
        right[2] = left[2]
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = grotto_keygen(BE, alpha, s0, s1, n, es)
    y0 = grotto_mic_eval(BE, masked, key, s0, 0, left, right, n, es).reshape(N, I)
    y1 = grotto_mic_eval(BE, masked, key, s1, 1, left, right, n, es).reshape(N, I)
    # signed n-bit interval containment of the unmasked x (tests/grotto/grotto_mic_eval_test_case.h)
    for i in range(N):
        xs = signed_nbit(int(x[i]), n)
        for q in range(I):
            want = int(signed_nbit(int(left[q]), n) <= xs <= signed_nbit(int(right[q]), n))
            assert (int(y0[i, q]) ^ int(y1[i, q])) & 1 == want, (i, q)
    lut = rand_elems(rng, 2 * I, 8 * es, es)
    e0, t0 = grotto_lut_eval(BE, masked, key, s0, 0, left, right, lut, n, 8 * es, es)
    e1, t1 = grotto_lut_eval(BE, masked, key, s1, 1, left, right, lut, n, 8 * es, es)
    W = mask(8 * es)
    for i in range(N):
        bits = [(int(y0[i, q]) ^ int(y1[i, q])) & 1 for q in range(I)]
        sh0 = [int(y0[i, q]) & 1 for q in range(I)]
        sh1 = [int(y1[i, q]) & 1 for q in range(I)]
        assert int(e0[i]) == sum(sh0) & W
        assert int(e1[i]) == (-sum(sh1)) & W
        for j in range(2):
            assert int(t0[i * 2 + j]) == sum(b * int(lut[q + j * I]) for q, b in enumerate(sh0)) & W
        del bits


def signed_full_partition(n, I):
    """The signed n-bit range cut into I contiguous intervals, as unsigned n-bit
    patterns (tests/common/nbit_signed.hpp:136-211): interval k covers offsets
    [2^n k / I, 2^n (k+1) / I - 1] from -2^(n-1)."""
    dom, umin, M = 1 << n, 1 << (n - 1), mask(n)
    left = [(umin + dom * k // I) & M for k in range(I)]
    right = [(umin + dom * (k + 1) // I - 1) & M for k in range(I)]
    return left, right


@pytest.mark.parametrize("n,m,es,I,lutNum,zero_alpha,gaps", [
    (6, 8, 1, 3, 2, True, False),     # the reference's grid (grotto_interval_lut_eval_cpu.cpp:141-163)
    (6, 16, 2, 3, 2, True, False),
    (6, 32, 4, 3, 2, True, False),
    (6, 64, 8, 3, 2, True, False),
    (8, 8, 1, 5, 3, False, False),
    (12, 16, 2, 7, 2, False, True),
    (20, 32, 4, 4, 2, False, True),
    (33, 64, 8, 6, 4, False, True),
])
def test_grotto_interval_lut_reconstruction(n, m, es, I, lutNum, zero_alpha, gaps):
    """E*T reconstructs LUT[matched] for both parties together: E = E0 + E1 is +1 or -1 where an
    interval matches and 0 where none does, and T = T0 + T1 == E * LUT[matched]. A swapped LUT
    index, a missing party sign or a dropped party-1 term fails one of these."""
    rng = np.random.default_rng(1000 * n + m)
    N = 600
    left, right = signed_full_partition(n, I)
    if gaps:  # drop every third interval: points in them match nothing
        keep = [k for k in range(I) if k % 3 != 1]
        left, right = [left[k] for k in keep], [right[k] for k in keep]
        I = len(left)
    left = np.array(left, dtype=DT[es])
    right = np.array(right, dtype=DT[es])
    alpha = np.zeros(N, dtype=DT[es]) if zero_alpha else rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    x[:I], x[I:2 * I] = left, right  # every endpoint
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    lut = rand_elems(rng, I * lutNum, m, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key = grotto_keygen(BE, alpha, s0, s1, n, es)
    E0, T0 = grotto_lut_eval(BE, masked, key, s0, 0, left, right, lut, n, m, es)
    E1, T1 = grotto_lut_eval(BE, masked, key, s1, 1, left, right, lut, n, m, es)
    W, Mo = mask(8 * es), mask(m)
    matched_some = unmatched_some = False
    for i in range(N):
        xs = signed_nbit(int(x[i]), n)
        hits = [k for k in range(I) if signed_nbit(int(left[k]), n) <= xs <= signed_nbit(int(right[k]), n)]
        assert len(hits) <= 1
        e = (int(E0[i]) + int(E1[i])) & W
        for j in range(lutNum):
            t = (int(T0[i * lutNum + j]) + int(T1[i * lutNum + j])) & Mo
            want = int(lut[hits[0] + j * I]) if hits else 0
            assert (e * t) & Mo == want, (i, j)
            assert t == (e * want) & Mo, (i, j)
        assert (e in (1, W)) if hits else e == 0, i
        matched_some |= bool(hits)
        unmatched_some |= not hits
    assert matched_some and (unmatched_some or not gaps)


def test_grotto_errors():
    z = np.zeros(4096, dtype=np.uint8)
    ks = 16 * 2 + 8 + 2
    assert ocall("grottoKeyGen", z, ks, z, 1, z, 16, z, 16, 5, 1, 1) == -10
    assert ocall("grottoKeyGen", z, ks, z, 1, z, 16, z, 16, 8, 1, 1) == 0
    assert ocall("grottoEval", z, 1, z, 1, z, ks, z, 16, True, 2, 8, 1, 1, None, 0) == -12
    assert ocall("grottoEval", z, 1, z, 1, z, ks, z, 16, True, 0, 8, 1, 1, z, 3) == -13
    assert ocall("grottoMICEval", z, 1, z, 1, z, ks, z, 16, 0, z, 2, z, 1, 8, 1, 1, None, 0) == -6


# -------------------------------------------------------------- DCF-MIC -----


@pytest.mark.parametrize("n,m,es", [(8, 8, 1), (16, 16, 2), (24, 32, 4), (31, 7, 4), (64, 64, 8)])
def test_mic_reconstruction(n, m, es):
    rng = np.random.default_rng(n * m)
    N, I = 300, 6
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    x[:I] = 0
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    left, right = _intervals(rng, I, n, es)
    x[I:2 * I] = left
    x[2 * I:3 * I] = right
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key, z = mic_keygen(BE, alpha, s0, s1, left, right, n, m, es)
    # z closed form (src/impl/mic.h:27-40)
    M = mask(n)
    for i in range(4):
        a = int(alpha[i])
        for q in range(I):
            L, R = int(left[q]), int(right[q])
            aP, aQ, aQp, qp = (L + a) & M, (R + a) & M, (R + 1 + a) & M, (R + 1) & M
            want = (aP > aQ) - (aP > L) + (aQp > qp) + (aQ == M)
            assert int(z[i * I + q]) == want & mask(8 * es)
    rng2 = np.random.default_rng(1)
    z0 = rand_elems(rng2, N * I, 8 * es, es)
    z1 = ((z.astype(np.uint64) - z0) & np.uint64(mask(8 * es))).astype(DT[es])
    y0 = mic_eval(BE, masked, key, z0, s0, 0, left, right, n, m, es).reshape(N, I)
    y1 = mic_eval(BE, masked, key, z1, s1, 1, left, right, n, m, es).reshape(N, I)
    rec = (y0.astype(np.uint64) + y1) & np.uint64(mask(m))
    for i in range(N):
        xs = signed_nbit(int(x[i]), n)
        for q in range(I):
            want = int(signed_nbit(int(left[q]), n) <= xs <= signed_nbit(int(right[q]), n))
            assert int(rec[i, q]) == want, (i, q)


# --------------------------------------------------------------- spline -----


def _spline_case(rng, I, d, n, m, es):
    """tests/spline/spline_eval_test_case.h:103-152: contiguous intervals covering [0, 2^n)."""
    M = mask(n)
    r = sorted(int(v) for v in rand_elems(rng, I, n, es))
    for i in range(I - 1):
        lo = 0 if i == 0 else r[i - 1] + 1
        hi = M - (I - 1 - i)
        r[i] = lo + r[i] % (hi - lo + 1)
    r[-1] = M
    left = [0] + [v + 1 for v in r[:-1]]
    coeff = rand_elems(rng, I * (d + 1), m, es)
    return np.array(left, dtype=DT[es]), np.array(r, dtype=DT[es]), coeff


@pytest.mark.parametrize("n,m,es,I,d", [(8, 8, 1, 4, 2), (16, 16, 2, 5, 1), (20, 20, 4, 3, 3), (32, 32, 4, 16, 2),
                                        (48, 48, 8, 2, 2)])
def test_spline_reconstruction(n, m, es, I, d):
    rng = np.random.default_rng(n + I)
    N = 64
    left, right, coeff = _spline_case(rng, I, d, n, m, es)
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    x[:I] = left
    x[I:2 * I] = right
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key, e, beta = spline_keygen(BE, alpha, s0, s1, coeff, d, left, right, n, m, es)
    W, Mo = mask(8 * es), mask(m)
    g = I * (d + 1)
    # beta_i are the coefficients of f_i(y - alpha): check at random points
    for i in range(3):
        a = int(alpha[i])
        for q in range(I):
            c = [int(v) for v in coeff[q * (d + 1):(q + 1) * (d + 1)]]
            b = [int(v) for v in beta[i * g + q * (d + 1):i * g + (q + 1) * (d + 1)]]
            for y in (0, 1, 12345, Mo):
                lhs = sum(ck * (y - a) ** k for k, ck in enumerate(c)) & Mo
                rhs = sum(bk * y ** k for k, bk in enumerate(b)) & Mo
                assert lhs == rhs
    # split beta/e into random additive shares as a dealer would
    b0 = rand_elems(rng, len(beta), 8 * es, es)
    e0 = rand_elems(rng, len(e), 8 * es, es)
    b1 = ((beta.astype(np.uint64) - b0) & np.uint64(W)).astype(DT[es])
    e1 = ((e.astype(np.uint64) - e0) & np.uint64(W)).astype(DT[es])
    y0 = spline_eval(BE, masked, key, e0, b0, s0, 0, left, right, d, n, m, es)
    y1 = spline_eval(BE, masked, key, e1, b1, s1, 1, left, right, d, n, m, es)
    rec = (y0.astype(np.uint64) + y1) & np.uint64(Mo)
    for i in range(N):
        xv = int(x[i])
        q = next(k for k in range(I) if int(right[k]) >= xv)
        c = [int(v) for v in coeff[q * (d + 1):(q + 1) * (d + 1)]]
        want = sum(ck * xv ** k for k, ck in enumerate(c)) & Mo
        assert int(rec[i]) == want, i


def _unsigned_intervals_with_gap(rng, I, n, es):
    """Contiguous unsigned cover of [0, 2^n) (as _spline_case), then interval 1 removed, so
    points inside it match no interval and the next interval's left end is not reused."""
    left, right, _ = _spline_case(rng, I + 1, 0, n, n, es)
    keep = [k for k in range(I + 1) if k != 1]
    return left[keep].copy(), right[keep].copy()


@pytest.mark.parametrize("n,m,es,I,d,gap", [(24, 32, 4, 16, 2, False), (24, 32, 4, 6, 2, True),
                                            (12, 16, 2, 5, 2, False), (8, 16, 2, 4, 3, True),
                                            (12, 10, 2, 5, 2, False), (20, 64, 8, 3, 2, True)])
def test_spline_in_ne_out(n, m, es, I, d, gap):
    """bitWidthIn != bitWidthOut: the coefficients are shifted by alpha mod 2^out and evaluated at
    X = maskedX mod 2^in, so the two parties reconstruct f_q(X - alpha) mod 2^out at the integer
    X - alpha (negative when the mask wrapped, X < alpha), q being the interval that holds the
    unmasked x = (X - alpha) mod 2^in, and 0 when no interval holds it. Shifting mod 2^in, or
    evaluating at x instead of X, fails this for X < alpha when n < m."""
    rng = np.random.default_rng(7 * n + m + I)
    N = 256
    if gap:
        left, right = _unsigned_intervals_with_gap(rng, I, n, es)
    else:
        left, right, _ = _spline_case(rng, I, d, n, m, es)
    I = len(left)
    coeff = rand_elems(rng, I * (d + 1), m, es)
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    x[:I], x[I:2 * I] = left, right
    M = mask(n)
    X = ((x.astype(np.uint64) + alpha) & np.uint64(M)).astype(DT[es])
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    key, e, beta = spline_keygen(BE, alpha, s0, s1, coeff, d, left, right, n, m, es)
    W, Mo = mask(8 * es), mask(m)
    b0 = rand_elems(rng, len(beta), 8 * es, es)
    e0 = rand_elems(rng, len(e), 8 * es, es)
    b1 = ((beta.astype(np.uint64) - b0) & np.uint64(W)).astype(DT[es])
    e1 = ((e.astype(np.uint64) - e0) & np.uint64(W)).astype(DT[es])
    y0 = spline_eval(BE, X, key, e0, b0, s0, 0, left, right, d, n, m, es)
    y1 = spline_eval(BE, X, key, e1, b1, s1, 1, left, right, d, n, m, es)
    wrapped = 0
    for i in range(N):
        Xi, a = int(X[i]), int(alpha[i])
        xi = (Xi - a) & M
        assert xi == int(x[i])
        hits = [k for k in range(I) if int(left[k]) <= xi <= int(right[k])]
        assert len(hits) <= 1
        if hits:
            c = [int(v) for v in coeff[hits[0] * (d + 1):(hits[0] + 1) * (d + 1)]]
            want = sum(ck * (Xi - a) ** k for k, ck in enumerate(c)) & Mo
        else:
            want = 0
        assert (int(y0[i]) + int(y1[i])) & Mo == want, (i, Xi, a)
        wrapped += Xi < a
    assert 0 < wrapped < N


def test_spline_cache_required():
    z = np.zeros(1 << 16, dtype=np.uint8)
    n, m, es, I, d, N = 8, 8, 1, 2, 1, 1
    g = I * (d + 1)
    ks = 16 * n + es * n * g + 2 * es + es * g
    args = [z, 1, z, 1, z, ks, z, g, z, g, z, 16, 0, z, I, z, I, I, d, n, m, es, N]
    assert ocall("dcfSplineEval", *args, None, 0) == -13
    cs = 16 * n + es * n * g + es * (2 * g + d + 1)
    assert ocall("dcfSplineEval", *args, z, cs) == 0


# ----------------------------------------------------------------- OTTT -----


@pytest.mark.parametrize("n,es", [(3, 1), (5, 1), (8, 2), (10, 4), (9, 8)])
@pytest.mark.parametrize("lutNum", [1, 3])
def test_ottt_reconstruction(n, es, lutNum):
    rng = np.random.default_rng(n * 7 + lutNum)
    N = 257
    alpha = rand_elems(rng, N, n, es)
    x = rand_elems(rng, N, n, es)
    masked = ((x.astype(np.uint64) + alpha) & np.uint64(mask(n))).astype(DT[es])
    lut = rand_elems(rng, lutNum << n, 8 * es, es)
    key0 = rng.integers(0, 256, size=N * (1 << n) // 8, dtype=np.uint8)
    key1 = ottt_keygen(BE, key0, alpha, n, es)
    diff = np.unpackbits(key0 ^ key1, bitorder="little").reshape(N, 1 << n)
    assert np.array_equal(np.argmax(diff, axis=1), alpha.astype(np.int64)) and diff.sum() == N
    E0, T0 = ottt_eval(BE, masked, key0, 0, lut, n, es)
    E1, T1 = ottt_eval(BE, masked, key1, 1, lut, n, es)
    W = np.uint64(mask(8 * es))
    E = (E0.astype(np.uint64) + E1) & W
    T = ((T0.astype(np.uint64) + T1) & W).reshape(N, lutNum)
    got = (E[:, None] * T) & W
    want = np.stack([lut[j * (1 << n) + x.astype(np.int64)] for j in range(lutNum)], axis=1).astype(np.uint64)
    assert np.array_equal(got, want)
    # E shares are (minus) popcounts of the key bits
    pc = np.unpackbits(key0, bitorder="little").reshape(N, -1).sum(axis=1).astype(np.uint64)
    assert np.array_equal(E0.astype(np.uint64), (np.uint64(0) - pc) & W)


def test_ottt_errors():
    z = np.zeros(64, dtype=np.uint8)
    assert ocall("otttKeyGen", z, 1, z, 1, 2, 1, 1) == -10
    assert ocall("otttLutEval", z, 1, z, 1, z, 1, z, 1, 0, z, 7, 3, 1, 1) == -5  # table size -> seed-size code
    assert ocall("otttLutEval", z, 1, z, 1, z, 1, z, 1, 3, z, 8, 3, 1, 1) == -12
