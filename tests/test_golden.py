"""The oracle reproduces the committed golden vectors, and the vectors satisfy
each primitive's reconstruction property on their own (so a fixture that was
written by a broken oracle would fail here too)."""
import numpy as np
import pytest

import fss_testlib as F
from golden_cases import REPLAYS, _cases, load

O = F.OracleBackend()


@pytest.mark.parametrize("name", sorted(REPLAYS))
def test_oracle_reproduces_golden(name):
    REPLAYS[name](O)


def test_prng_golden():
    z = load("prng")
    lib = F.oracle()
    g = lib.oracle_prngInit()
    assert lib.oracle_prngSetCurrentSeed(g, z["seed"].tobytes(), z["counter"].tobytes()) == 0
    out = np.zeros_like(z["out_u32"])
    assert lib.oracle_prngGen(g, out.ctypes.data, 32, 4, out.size) == 0
    lib.oracle_prngRelease(g)
    assert np.array_equal(out, z["out_u32"])
    # AES-CTR: block k = AES_seed(counter + k), 128-bit little-endian counter (src/cpu/prng.cpp:20-76)
    if F.Aes.available():
        ctr = int.from_bytes(z["counter"].tobytes(), "little")
        want = b"".join(F.Aes.encrypt(z["seed"].tobytes(), ((ctr + k) % 2**128).to_bytes(16, "little"))
                        for k in range(65))
        assert out.tobytes() == want[:out.nbytes]


def test_dcf_dpf_golden_reconstruct():
    for prim, pred in (("dcf", np.less), ("dpf", np.equal)):
        z = load(prim)
        for c in _cases(z):
            n, m, es, g, N = (int(v) for v in z[f"{c}_params"])
            rec = (z[f"{c}_y0"].astype(np.uint64) + z[f"{c}_y1"]) & np.uint64(F.mask(m))
            want = np.where(pred(z[f"{c}_x"], z[f"{c}_alpha"])[:, None], z[f"{c}_beta"].reshape(N, g), 0)
            assert np.array_equal(rec.reshape(N, g), want.astype(np.uint64) & np.uint64(F.mask(m))), (prim, c)


def test_grotto_golden_reconstruct():
    z = load("grotto")
    for c in _cases(z):
        x, a = z[f"{c}_x"], z[f"{c}_alpha"]
        for mode, want in (("lt", x < a), ("le", x <= a), ("eq", x == a)):
            assert np.array_equal((z[f"{c}_{mode}0"] ^ z[f"{c}_{mode}1"]) & 1, want.astype(x.dtype)), (c, mode)


def test_mic_ottt_golden_reconstruct():
    z = load("mic")
    n, m, es, I, N = (int(v) for v in z["params"])
    rec = ((z["y0"].astype(np.uint64) + z["y1"]) & np.uint64(0xFFFFFFFF)).reshape(N, I)
    v = (z["x"].astype(np.int64) - z["alpha"].astype(np.int64)) & F.mask(n)  # the fixture's x is masked
    for q in range(I):
        lo, hi = int(z["left"][q]), int(z["right"][q])
        want = [int(F.signed_nbit(lo, n) <= F.signed_nbit(int(t), n) <= F.signed_nbit(hi, n)) for t in v]
        if F.signed_nbit(lo, n) <= F.signed_nbit(hi, n):
            assert rec[:, q].tolist() == want, q
    z = load("ottt")
    n, es, N, lutNum = (int(v) for v in z["params"])
    W = np.uint64(0xFFFF)
    E = (z["e0"].astype(np.uint64) + z["e1"]) & W
    T = ((z["t0"].astype(np.uint64) + z["t1"]) & W).reshape(N, lutNum)
    x = (z["x"].astype(np.int64) - z["alpha"].astype(np.int64)) & F.mask(n)
    want = np.stack([z["lut"][j * (1 << n) + x] for j in range(lutNum)], axis=1).astype(np.uint64)
    assert np.array_equal((E[:, None] * T) & W, want)
