"""Pins the oracle's AES-128 and AES-CTR PRNG.

AES: FIPS-197 known-answer vectors (Appendix A.1 key expansion, Appendix B
cipher example, Appendix C.1), and agreement of the oracle's two encryptors
(spec-derived portable and AES-NI) with the system libcrypto on random inputs.

PRNG: the reference stream (src/cpu/prng.cpp:20-76) is block k = AES_seed(ctr+k)
with a 128-bit little-endian counter; checked against libcrypto, including a
carry out of the low counter word and the counter update after a call.
"""
import ctypes

import numpy as np
import pytest

from fss_testlib import Aes, le_block, oracle


def _enc(key: bytes, pt: bytes, portable: bool) -> bytes:
    lib = oracle()
    lib.oracle_aes_force_portable(1 if portable else 0)
    try:
        k = ctypes.create_string_buffer(key, 16)
        p = ctypes.create_string_buffer(pt, 16)
        o = ctypes.create_string_buffer(16)
        lib.oracle_aes128_encrypt_blocks(ctypes.addressof(k), ctypes.addressof(p), ctypes.addressof(o), 1)
        return o.raw
    finally:
        lib.oracle_aes_force_portable(0)


KATS = [
    # FIPS-197 Appendix C.1
    ("000102030405060708090a0b0c0d0e0f", "00112233445566778899aabbccddeeff", "69c4e0d86a7b0430d8cdb78070b4c55a"),
    # FIPS-197 Appendix B
    ("2b7e151628aed2a6abf7158809cf4f3c", "3243f6a8885a308d313198a2e0370734", "3925841d02dc09fbdc118597196a0b32"),
]


@pytest.mark.parametrize("portable", [True, False])
@pytest.mark.parametrize("key,pt,ct", KATS)
def test_fips197_kat(key, pt, ct, portable):
    assert _enc(bytes.fromhex(key), bytes.fromhex(pt), portable).hex() == ct


def test_key_expansion_appendix_a1():
    lib = oracle()
    k = ctypes.create_string_buffer(bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c"), 16)
    rk = ctypes.create_string_buffer(176)
    lib.oracle_aes128_expand_portable(ctypes.addressof(k), ctypes.addressof(rk))
    # w[4..7] and w[40..43] from FIPS-197 Appendix A.1
    assert rk.raw[16:32].hex() == "a0fafe1788542cb123a339392a6c7605"
    assert rk.raw[160:176].hex() == "d014f9a8c9ee2589e13f0cc8b6630ca6"


def test_sbox_known_points():
    lib = oracle()
    sbox = [lib.oracle_aes_sbox(i) for i in range(256)]
    assert sorted(sbox) == list(range(256))
    assert sbox[0x00] == 0x63 and sbox[0x53] == 0xED and sbox[0xFF] == 0x16


@pytest.mark.skipif(not Aes.available(), reason="system libcrypto not found")
def test_random_vs_libcrypto():
    rng = np.random.default_rng(7)
    for _ in range(300):
        key = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
        pt = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
        want = Aes.encrypt(key, pt)
        assert _enc(key, pt, True) == want
        assert _enc(key, pt, False) == want


def test_multiblock_matches_single():
    lib = oracle()
    rng = np.random.default_rng(3)
    key = rng.integers(0, 256, 16, dtype=np.uint8)
    pts = rng.integers(0, 256, 16 * 9, dtype=np.uint8)
    out = np.zeros_like(pts)
    lib.oracle_aes128_encrypt_blocks(key.ctypes.data, pts.ctypes.data, out.ctypes.data, 9)
    for b in range(9):
        assert out[16 * b:16 * b + 16].tobytes() == _enc(key.tobytes(), pts[16 * b:16 * b + 16].tobytes(), True)


def _prng_stream(seed: bytes, ctr_lo: int, ctr_hi: int, nbytes: int) -> bytes:
    out = b""
    k = 0
    while len(out) < nbytes:
        lo = ctr_lo + k
        hi = ctr_hi + (lo >> 64)
        out += Aes.encrypt(seed, le_block(lo, hi))
        k += 1
    return out[:nbytes]


@pytest.mark.skipif(not Aes.available(), reason="system libcrypto not found")
@pytest.mark.parametrize("es,num", [(1, 37), (2, 1000), (4, 5), (8, 3), (4, 4097)])
def test_prng_closed_form(es, num):
    lib = oracle()
    seed = bytes(range(16))
    ctr_lo = 2**64 - 2  # the low word overflows inside the call
    ctr = le_block(ctr_lo, 5)
    g = lib.oracle_prngInit()
    assert lib.oracle_prngSetCurrentSeed(g, seed, ctr) == 0
    dst = np.zeros(es * num, dtype=np.uint8)
    assert lib.oracle_prngGen(g, dst.ctypes.data, 8 * es, es, num) == 0
    assert dst.tobytes() == _prng_stream(seed, ctr_lo, 5, es * num)
    s2 = ctypes.create_string_buffer(16)
    c2 = ctypes.create_string_buffer(16)
    lib.oracle_prngGetCurrentSeed(g, s2, c2)
    blocks = (es * num + 15) // 16
    lo = ctr_lo + blocks
    assert c2.raw == le_block(lo, 5 + (lo >> 64))
    assert s2.raw == seed
    lib.oracle_prngRelease(g)


def test_prng_errors():
    lib = oracle()
    g = lib.oracle_prngInit()
    dst = np.zeros(16, dtype=np.uint8)
    assert lib.oracle_prngGen(None, dst.ctypes.data, 8, 1, 16) == -17
    assert lib.oracle_prngGen(g, None, 8, 1, 16) == -17
    assert lib.oracle_prngGen(g, dst.ctypes.data, 0, 1, 16) == -18
    assert lib.oracle_prngGen(g, dst.ctypes.data, 9, 1, 16) == -18
    assert lib.oracle_prngSetCurrentSeed(g, None, None) == -17
    lib.oracle_prngRelease(g)
