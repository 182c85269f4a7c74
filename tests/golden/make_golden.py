"""Writes the golden vectors under tests/golden/ from the CPU oracle.

The reference ships no known-answer vectors for its FSS primitives and its CPU
path cannot be built here, so these fixtures are produced by the oracle (C
restatement of the reference CPU path), which tests/test_oracle_*.py pins
against FIPS-197, libcrypto and the closed forms of each primitive.  The
fixtures freeze those pinned outputs: tests/test_golden.py checks that the
oracle still reproduces them, tests/test_gpu_golden.py that the device library
does.  Every input comes from numpy's default_rng with a fixed seed.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
import zlib
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

import fss_testlib as F  # noqa: E402

O = F.OracleBackend()

DCF_CASES = [  # name, n, m, es, g, N
    ("n16_u16", 16, 16, 2, 1, 64),  # the reference's CPU-sized configuration, shrunk
    ("n32_u32", 32, 32, 4, 1, 64),
    ("n64_u64_g3", 64, 64, 8, 3, 16),
    ("n24_m20_g6", 24, 20, 4, 6, 16),  # value conversion through the AES refill
    ("n8_u8", 8, 8, 1, 1, 32),
]
DPF_CASES = [("n32_u32", 32, 32, 4, 1, 64), ("n16_m8", 16, 8, 2, 1, 32), ("n64_g2", 64, 64, 8, 2, 16)]
GROTTO_CASES = [("n16_u16", 16, 2, 64), ("n32_u32", 32, 4, 32)]


def dcf(z):
    for name, n, m, es, g, N in DCF_CASES:
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        alpha, x = F.rand_elems(rng, N, n, es), F.rand_elems(rng, N, n, es)
        beta = F.rand_elems(rng, N * g, m, es)
        s0, s1 = F.rand_seeds(rng, N), F.rand_seeds(rng, N)
        key = F.dcf_keygen(O, alpha, beta, s0, s1, n, m, g, es)
        z.update({f"{name}_params": np.array([n, m, es, g, N]), f"{name}_alpha": alpha, f"{name}_beta": beta,
                  f"{name}_x": x, f"{name}_seed0": s0, f"{name}_seed1": s1, f"{name}_key": key,
                  f"{name}_y0": F.dcf_eval(O, x, key, s0, 0, n, m, g, es),
                  f"{name}_y1": F.dcf_eval(O, x, key, s1, 1, n, m, g, es)})


def dpf(z):
    for name, n, m, es, g, N in DPF_CASES:
        rng = np.random.default_rng(zlib.crc32(name.encode()) + 1)
        alpha, x = F.rand_elems(rng, N, n, es), F.rand_elems(rng, N, n, es)
        x[::3] = alpha[::3]
        beta = F.rand_elems(rng, N * g, m, es)
        s0, s1 = F.rand_seeds(rng, N), F.rand_seeds(rng, N)
        key = F.dpf_keygen(O, alpha, beta, s0, s1, n, m, g, es)
        z.update({f"{name}_params": np.array([n, m, es, g, N]), f"{name}_alpha": alpha, f"{name}_beta": beta,
                  f"{name}_x": x, f"{name}_seed0": s0, f"{name}_seed1": s1, f"{name}_key": key,
                  f"{name}_y0": F.dpf_eval(O, x, key, s0, 0, n, m, g, es),
                  f"{name}_y1": F.dpf_eval(O, x, key, s1, 1, n, m, g, es)})


def grotto(z):
    for name, n, es, N in GROTTO_CASES:
        rng = np.random.default_rng(zlib.crc32(name.encode()) + 2)
        alpha, x = F.rand_elems(rng, N, n, es), F.rand_elems(rng, N, n, es)
        x[::4] = alpha[::4]
        x[1::5] = 0
        s0, s1 = F.rand_seeds(rng, N), F.rand_seeds(rng, N)
        key = F.grotto_keygen(O, alpha, s0, s1, n, es)
        z.update({f"{name}_params": np.array([n, es, N]), f"{name}_alpha": alpha, f"{name}_x": x,
                  f"{name}_seed0": s0, f"{name}_seed1": s1, f"{name}_key": key})
        for party, sd in ((0, s0), (1, s1)):
            z[f"{name}_lt{party}"] = F.grotto_eval(O, x, key, sd, party, n, es, False)
            z[f"{name}_le{party}"] = F.grotto_eval(O, x, key, sd, party, n, es, True)
            z[f"{name}_eq{party}"] = F.grotto_eq_eval(O, x, key, sd, party, n, es)


def mic(z):
    n, m, es, I, N = 24, 32, 4, 4, 32
    rng = np.random.default_rng(24)
    alpha, x = F.rand_elems(rng, N, n, es), F.rand_elems(rng, N, n, es)
    left = np.array([0, 1 << 20, 5 << 20, 9 << 20], dtype=np.uint32)
    right = np.array([(1 << 20) - 1, (5 << 20) - 1, (9 << 20) - 1, (1 << 24) - 1], dtype=np.uint32)
    s0, s1 = F.rand_seeds(rng, N), F.rand_seeds(rng, N)
    key, zz = F.mic_keygen(O, alpha, s0, s1, left, right, n, m, es)
    z0 = F.rand_elems(rng, N * I, 32, es)
    z1 = ((zz.astype(np.uint64) - z0) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    z.update({"params": np.array([n, m, es, I, N]), "alpha": alpha, "x": x, "left": left, "right": right,
              "seed0": s0, "seed1": s1, "key": key, "z": zz, "z0": z0, "z1": z1,
              "y0": F.mic_eval(O, x, key, z0, s0, 0, left, right, n, m, es),
              "y1": F.mic_eval(O, x, key, z1, s1, 1, left, right, n, m, es)})


def spline(z):
    n, m, es, I, d, N = 16, 16, 2, 3, 2, 16
    rng = np.random.default_rng(16)
    alpha, x = F.rand_elems(rng, N, n, es), F.rand_elems(rng, N, n, es)
    left = np.array([0, 20000, 41000], dtype=np.uint16)
    right = np.array([19999, 40999, 65535], dtype=np.uint16)
    coeff = F.rand_elems(rng, I * (d + 1), m, es)
    s0, s1 = F.rand_seeds(rng, N), F.rand_seeds(rng, N)
    key, e, beta = F.spline_keygen(O, alpha, s0, s1, coeff, d, left, right, n, m, es)
    g = I * (d + 1)
    e0, b0 = F.rand_elems(rng, N * g, 16, es), F.rand_elems(rng, N * g, 16, es)
    e1 = ((e.astype(np.int64) - e0) & 0xFFFF).astype(np.uint16)
    b1 = ((beta.astype(np.int64) - b0) & 0xFFFF).astype(np.uint16)
    z.update({"params": np.array([n, m, es, I, d, N]), "alpha": alpha, "x": x, "left": left, "right": right,
              "coeff": coeff, "seed0": s0, "seed1": s1, "key": key, "e": e, "beta": beta, "e0": e0, "b0": b0,
              "e1": e1, "b1": b1,
              "y0": F.spline_eval(O, x, key, e0, b0, s0, 0, left, right, d, n, m, es),
              "y1": F.spline_eval(O, x, key, e1, b1, s1, 1, left, right, d, n, m, es)})


def ottt(z):
    n, es, N, lutNum = 8, 2, 16, 2
    rng = np.random.default_rng(8)
    alpha, x = F.rand_elems(rng, N, n, es), F.rand_elems(rng, N, n, es)
    key0 = rng.integers(0, 256, size=N * (1 << n) // 8, dtype=np.uint8)
    key1 = F.ottt_keygen(O, key0, alpha, n, es)
    lut = F.rand_elems(rng, lutNum << n, 16, es)
    e0, t0 = F.ottt_eval(O, x, key0, 0, lut, n, es)
    e1, t1 = F.ottt_eval(O, x, key1, 1, lut, n, es)
    z.update({"params": np.array([n, es, N, lutNum]), "alpha": alpha, "x": x, "key0": key0, "key1": key1,
              "lut": lut, "e0": e0, "t0": t0, "e1": e1, "t1": t1})


def prng(z):
    lib = F.oracle()
    seed, ctr = bytes(range(16)), bytes(range(100, 116))
    g = lib.oracle_prngInit()
    assert lib.oracle_prngSetCurrentSeed(g, seed, ctr) == 0
    out = np.zeros(257, dtype=np.uint32)
    assert lib.oracle_prngGen(g, out.ctypes.data, 32, 4, 257) == 0
    lib.oracle_prngRelease(g)
    z.update({"seed": np.frombuffer(seed, np.uint8), "counter": np.frombuffer(ctr, np.uint8), "out_u32": out})


def main():
    for name, fn in (("dcf", dcf), ("dpf", dpf), ("grotto", grotto), ("mic", mic), ("spline", spline),
                     ("ottt", ottt), ("prng", prng)):
        z = {}
        fn(z)
        np.savez(HERE / f"{name}.npz", **z)
        print(name, sum(v.nbytes for v in z.values()), "bytes")


if __name__ == "__main__":
    main()
