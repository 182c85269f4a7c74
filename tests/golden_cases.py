"""Replays the golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) through a backend (oracle or device library) and
compares every key byte and share with the stored vectors."""
from __future__ import annotations

from pathlib import Path

import numpy as np

import fss_testlib as F

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _cases(z):
    return sorted({k.rsplit("_", 1)[0] for k in z if k.endswith("_params")})


def _eq(be, got, want, what):
    assert np.array_equal(be.get(got), want), what


def replay_dcf(be):
    z = load("dcf")
    for c in _cases(z):
        n, m, es, g, N = (int(v) for v in z[f"{c}_params"])
        key = F.dcf_keygen(be, z[f"{c}_alpha"], z[f"{c}_beta"], z[f"{c}_seed0"], z[f"{c}_seed1"], n, m, g, es)
        _eq(be, key, z[f"{c}_key"], (c, "key"))
        for p in (0, 1):
            _eq(be, F.dcf_eval(be, z[f"{c}_x"], key, z[f"{c}_seed{p}"], p, n, m, g, es), z[f"{c}_y{p}"], (c, p))


def replay_dpf(be):
    z = load("dpf")
    for c in _cases(z):
        n, m, es, g, N = (int(v) for v in z[f"{c}_params"])
        key = F.dpf_keygen(be, z[f"{c}_alpha"], z[f"{c}_beta"], z[f"{c}_seed0"], z[f"{c}_seed1"], n, m, g, es)
        _eq(be, key, z[f"{c}_key"], (c, "key"))
        for p in (0, 1):
            _eq(be, F.dpf_eval(be, z[f"{c}_x"], key, z[f"{c}_seed{p}"], p, n, m, g, es), z[f"{c}_y{p}"], (c, p))


def replay_grotto(be):
    z = load("grotto")
    for c in _cases(z):
        n, es, N = (int(v) for v in z[f"{c}_params"])
        key = F.grotto_keygen(be, z[f"{c}_alpha"], z[f"{c}_seed0"], z[f"{c}_seed1"], n, es)
        _eq(be, key, z[f"{c}_key"], (c, "key"))
        x = z[f"{c}_x"]
        for p in (0, 1):
            sd = z[f"{c}_seed{p}"]
            _eq(be, F.grotto_eval(be, x, key, sd, p, n, es, False), z[f"{c}_lt{p}"], (c, "lt", p))
            _eq(be, F.grotto_eval(be, x, key, sd, p, n, es, True), z[f"{c}_le{p}"], (c, "le", p))
            _eq(be, F.grotto_eq_eval(be, x, key, sd, p, n, es), z[f"{c}_eq{p}"], (c, "eq", p))


def replay_mic(be):
    z = load("mic")
    n, m, es, I, N = (int(v) for v in z["params"])
    key, zz = F.mic_keygen(be, z["alpha"], z["seed0"], z["seed1"], z["left"], z["right"], n, m, es)
    _eq(be, key, z["key"], "key")
    _eq(be, zz, z["z"], "z")
    for p in (0, 1):
        y = F.mic_eval(be, z["x"], key, z[f"z{p}"], z[f"seed{p}"], p, z["left"], z["right"], n, m, es)
        _eq(be, y, z[f"y{p}"], p)


def replay_spline(be):
    z = load("spline")
    n, m, es, I, d, N = (int(v) for v in z["params"])
    key, e, beta = F.spline_keygen(be, z["alpha"], z["seed0"], z["seed1"], z["coeff"], d, z["left"], z["right"], n,
                                   m, es)
    _eq(be, key, z["key"], "key")
    _eq(be, e, z["e"], "e")
    _eq(be, beta, z["beta"], "beta")
    for p in (0, 1):
        y = F.spline_eval(be, z["x"], key, z[f"e{p}"], z[f"b{p}"], z[f"seed{p}"], p, z["left"], z["right"], d, n, m,
                          es)
        _eq(be, y, z[f"y{p}"], p)


def replay_ottt(be):
    z = load("ottt")
    n, es, N, lutNum = (int(v) for v in z["params"])
    key1 = F.ottt_keygen(be, z["key0"], z["alpha"], n, es)
    _eq(be, key1, z["key1"], "key1")
    for p, k in ((0, be.put(z["key0"])), (1, key1)):
        e, t = F.ottt_eval(be, z["x"], k, p, z["lut"], n, es)
        _eq(be, e, z[f"e{p}"], ("E", p))
        _eq(be, t, z[f"t{p}"], ("T", p))


REPLAYS = {"dcf": replay_dcf, "dpf": replay_dpf, "grotto": replay_grotto, "mic": replay_mic,
           "spline": replay_spline, "ottt": replay_ottt}
