"""Shared test helpers.

* ``oracle()`` loads the CPU oracle (``oracle/liboracle.so``), the plain-C
  restatement of the reference CPU path used as the checker.
* ``device()`` loads the device library's C ABI (``FastFss_*`` symbols).
* ``Aes`` is an independent AES-128 from the system libcrypto, used to pin the
  oracle's AES and PRG conventions without trusting the oracle's own AES.
* ``SIGS`` lists the argument kinds of every entry point of the reference C API
  (include/FastFss/cpu/<prim>.h, include/FastFss/cuda/<prim>.h) so one
  argument list drives both libraries.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
ORACLE_DIR = REPO / "oracle"
ORACLE_SO = Path(os.environ.get("FSS_ORACLE_SO", str(ORACLE_DIR / "liboracle.so")))

# argument kinds: p = pointer, s = size_t, i = int, b = bool
SIGS = {
    "dcfKeyGen": "pspspspsps" + "sssss",
    "dcfEval": "pspspsps" + "i" + "sssss" + "ps",
    "dpfKeyGen": "pspspspsps" + "sssss",
    "dpfEval": "pspspsps" + "i" + "sssss" + "ps",
    "dpfEvalAll": "pspspsps" + "i" + "sssss" + "ps",
    "dpfEvalMulti": "pspspsps" + "i" + "ps" + "sssss" + "ps",
    "grottoKeyGen": "pspspsps" + "sss",
    "grottoEqEval": "pspspsps" + "i" + "sss" + "ps",
    "grottoEval": "pspspsps" + "b" + "i" + "sss" + "ps",
    "grottoMICEval": "pspspsps" + "i" + "psps" + "sss" + "ps",
    "grottoIntervalLutEval": "pspspspsps" + "i" + "pspsps" + "ssss" + "ps",
    "dcfMICKeyGen": "pspspspspspsps" + "ssss",
    "dcfMICEval": "pspspspsps" + "i" + "psps" + "ssss" + "ps",
    "dcfSplineKeyGen": "pspspspspspsps" + "s" + "psps" + "sssss",
    "dcfSplineEval": "pspspspspsps" + "i" + "psps" + "ssssss" + "ps",
    "otttKeyGen": "psps" + "sss",
    "otttLutEval": "pspspsps" + "i" + "ps" + "sss",
}

SIZE_SIGS = {
    "dcfGetKeyDataSize": "psssss",
    "dcfGetCacheDataSize": "psssss",
    "dpfGetKeyDataSize": "psssss",
    "dpfGetCacheDataSize": "psss",
    "grottoGetKeyDataSize": "psss",
    "grottoGetCacheDataSize": "psss",
    "dcfMICGetKeyDataSize": "pssss",
    "dcfMICGetCacheDataSize": "pssss",
    "dcfSplineGetKeyDataSize": "pssssss",
    "dcfSplineGetCacheDataSize": "pssssss",
    "otttGetKeyDataSize": "pss",
}

_KIND = {"p": ctypes.c_void_p, "s": ctypes.c_size_t, "i": ctypes.c_int, "b": ctypes.c_bool}


def _bind(lib, name, kinds, restype=ctypes.c_int):
    fn = getattr(lib, name)
    fn.argtypes = [_KIND[k] for k in kinds]
    fn.restype = restype
    return fn


def ptr(a):
    """Address of a numpy array / torch tensor / None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data if a.size else None
    if hasattr(a, "data_ptr"):
        return a.data_ptr() if a.numel() else None
    return a


def nbytes(a):
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.nbytes
    return a.numel() * a.element_size()


_oracle = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def oracle():
    """The CPU oracle with `name(...)` wrappers taking the reference argument lists."""
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            build_oracle()
        lib = ctypes.CDLL(str(ORACLE_SO))
        for name, kinds in SIGS.items():
            _bind(lib, "oracle_" + name, kinds)
        for name, kinds in SIZE_SIGS.items():
            _bind(lib, "oracle_" + name, kinds)
        _bind(lib, "oracle_prngInit", "", ctypes.c_void_p)
        _bind(lib, "oracle_prngRelease", "p", None)
        _bind(lib, "oracle_prngSetCurrentSeed", "ppp")
        _bind(lib, "oracle_prngGetCurrentSeed", "ppp")
        _bind(lib, "oracle_prngGen", "ppsss")
        _bind(lib, "oracle_aes128_encrypt_blocks", "ppps", None)
        _bind(lib, "oracle_aes128_expand_portable", "pp", None)
        _bind(lib, "oracle_aes128_encrypt_portable", "ppp", None)
        _bind(lib, "oracle_aes_force_portable", "i", None)
        _bind(lib, "oracle_aes_uses_aesni", "")
        _bind(lib, "oracle_aes_sbox", "i", ctypes.c_uint8)
        _bind(lib, "oracle_setNumThreads", "i")
        _bind(lib, "oracle_getNumThreads", "")
        _oracle = lib
    return _oracle


def ocall(name, *args):
    """Call oracle_<name> with numpy arrays / ints, returning the error code."""
    fn = getattr(oracle(), "oracle_" + name)
    return fn(*[ptr(a) if (isinstance(a, np.ndarray) or a is None) else a for a in args])


# ---------------------------------------------------------------------------
# independent AES-128 (system libcrypto)
# ---------------------------------------------------------------------------


class Aes:
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            name = ctypes.util.find_library("crypto")
            if name is None:
                return None
            lib = ctypes.CDLL(name)
            lib.AES_set_encrypt_key.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
            lib.AES_set_encrypt_key.restype = ctypes.c_int
            lib.AES_encrypt.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
            lib.AES_encrypt.restype = None
            cls._lib = lib
        return cls._lib

    @classmethod
    def available(cls):
        try:
            return cls.lib() is not None
        except (OSError, AttributeError):
            return False

    @classmethod
    def encrypt(cls, key: bytes, block: bytes) -> bytes:
        lib = cls.lib()
        ks = ctypes.create_string_buffer(512)
        assert lib.AES_set_encrypt_key(bytes(key), 128, ks) == 0
        out = ctypes.create_string_buffer(16)
        lib.AES_encrypt(bytes(block), out, ks)
        return out.raw


def le_block(lo: int, hi: int) -> bytes:
    return (lo & (2**64 - 1)).to_bytes(8, "little") + (hi & (2**64 - 1)).to_bytes(8, "little")


def clear_lsb(b: bytes) -> bytes:
    return bytes([b[0] & 0xFE]) + b[1:]


def xor16(a: bytes, b: bytes) -> bytes:
    return bytes(x ^ y for x, y in zip(a, b))


# ---------------------------------------------------------------------------
# input generation
# ---------------------------------------------------------------------------

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def mask(bits: int) -> int:
    return (1 << bits) - 1


def rand_elems(rng, n, bits, es):
    """n uniform values of `bits` bits stored in es-byte elements."""
    raw = rng.integers(0, 2**63, size=n, dtype=np.uint64) ^ (
        rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63)
    )
    if bits < 64:
        raw &= np.uint64(mask(bits))
    return raw.astype(DT[es])


def rand_seeds(rng, n):
    return rng.integers(0, 256, size=16 * n, dtype=np.uint8)


def elem_to_int(a):
    return [int(v) for v in a]


# ---------------------------------------------------------------------------
# backends: the oracle on host memory, the device library on HBM
# ---------------------------------------------------------------------------


class OracleBackend:
    """Runs the reference argument lists on the CPU oracle (numpy buffers)."""

    name = "oracle"

    def put(self, a):
        return None if a is None else np.ascontiguousarray(a)

    def empty(self, nbytes, dtype=np.uint8):
        return np.zeros(nbytes // np.dtype(dtype).itemsize, dtype=dtype)

    def get(self, a):
        return a

    def call(self, name, *args):
        return ocall(name, *args)


class HostBackend(OracleBackend):
    """Runs the reference argument lists on numpy (host) buffers through the device library's
    host-buffer entry points (FastFss_cpu_*), which evaluate on the GPU."""

    name = "host-api"

    def __init__(self):
        from fastfunctionsecretsharing_amd import _lib

        self.lib = _lib.load()

    def call(self, name, *args):
        prefix = "FastFss_" if name.startswith(("dcfGet", "dpfGet", "grottoGet", "dcfMICGet", "dcfSplineGet",
                                                  "otttGet")) else "FastFss_cpu_"
        return getattr(self.lib, prefix + name)(*[ptr(a) if (isinstance(a, np.ndarray) or a is None) else a
                                                  for a in args])


def _size(be_call, name, *args):
    out = ctypes.c_size_t(0)
    rc = be_call(name, ctypes.addressof(out), *args)
    assert rc == 0, f"{name} -> {rc}"
    return out.value


def key_size(be, name, *args):
    return _size(be.call, name, *args)


def run(be, name, *args):
    rc = be.call(name, *args)
    assert rc == 0, f"{be.name}.{name} failed: {rc}"


def dcf_keygen(be, alpha, beta, s0, s1, n, m, g, es):
    N = len(alpha)
    ks = key_size(be, "dcfGetKeyDataSize", n, m, g, es, N)
    key = be.empty(ks)
    a, b, d0, d1 = be.put(alpha), be.put(beta), be.put(s0), be.put(s1)
    run(be, "dcfKeyGen", key, ks, a, nbytes(a), b, nbytes(b), d0, nbytes(d0), d1, nbytes(d1), n, m, g, es, N)
    return key


def dcf_eval(be, x, key, seed, party, n, m, g, es):
    N = len(x)
    out = be.empty(N * es * g, DT[es])
    xx, sd = be.put(x), be.put(seed)
    run(be, "dcfEval", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), sd, nbytes(sd), party, n, m, g, es, N,
        None, 0)
    return out


def dpf_keygen(be, alpha, beta, s0, s1, n, m, g, es):
    N = len(alpha)
    ks = key_size(be, "dpfGetKeyDataSize", n, m, g, es, N)
    key = be.empty(ks)
    a, b, d0, d1 = be.put(alpha), be.put(beta), be.put(s0), be.put(s1)
    run(be, "dpfKeyGen", key, ks, a, nbytes(a), b, nbytes(b), d0, nbytes(d0), d1, nbytes(d1), n, m, g, es, N)
    return key


def dpf_eval(be, x, key, seed, party, n, m, g, es):
    N = len(x)
    out = be.empty(N * es * g, DT[es])
    xx, sd = be.put(x), be.put(seed)
    run(be, "dpfEval", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), sd, nbytes(sd), party, n, m, g, es, N,
        None, 0)
    return out


def dpf_eval_all(be, x, key, seed, party, n, m, g, es):
    N = len(x)
    out = be.empty(N * es * g * (1 << n), DT[es])
    xx, sd = be.put(x), be.put(seed)
    run(be, "dpfEvalAll", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), sd, nbytes(sd), party, n, m, g, es,
        N, None, 0)
    return out


def dpf_eval_multi(be, x, key, seed, party, points, n, m, g, es):
    N = len(x)
    P = len(points)
    out = be.empty(N * es * g * P, DT[es])
    xx, sd, pt = be.put(x), be.put(seed), be.put(points)
    run(be, "dpfEvalMulti", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), sd, nbytes(sd), party, pt,
        nbytes(pt), n, m, g, es, N, None, 0)
    return out


def grotto_keygen(be, alpha, s0, s1, n, es):
    N = len(alpha)
    ks = key_size(be, "grottoGetKeyDataSize", n, es, N)
    key = be.empty(ks)
    a, d0, d1 = be.put(alpha), be.put(s0), be.put(s1)
    run(be, "grottoKeyGen", key, ks, a, nbytes(a), d0, nbytes(d0), d1, nbytes(d1), n, es, N)
    return key


def grotto_eval(be, x, key, seed, party, n, es, equal_bound):
    N = len(x)
    out = be.empty(N * es, DT[es])
    xx, sd = be.put(x), be.put(seed)
    run(be, "grottoEval", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), sd, nbytes(sd), bool(equal_bound),
        party, n, es, N, None, 0)
    return out


def grotto_eq_eval(be, x, key, seed, party, n, es):
    N = len(x)
    out = be.empty(N * es, DT[es])
    xx, sd = be.put(x), be.put(seed)
    run(be, "grottoEqEval", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), sd, nbytes(sd), party, n, es, N,
        None, 0)
    return out


def grotto_mic_eval(be, x, key, seed, party, left, right, n, es):
    N = len(x)
    I = len(left)
    out = be.empty(N * es * I, DT[es])
    xx, sd, lf, rt = be.put(x), be.put(seed), be.put(left), be.put(right)
    run(be, "grottoMICEval", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), sd, nbytes(sd), party, lf,
        nbytes(lf), rt, nbytes(rt), n, es, N, None, 0)
    return out


def grotto_lut_eval(be, x, key, seed, party, left, right, lut, n, m, es):
    N = len(x)
    I = len(left)
    lutNum = len(lut) // I
    oe = be.empty(N * es, DT[es])
    ot = be.empty(N * es * lutNum, DT[es])
    xx, sd, lf, rt, lt = be.put(x), be.put(seed), be.put(left), be.put(right), be.put(lut)
    run(be, "grottoIntervalLutEval", oe, nbytes(oe), ot, nbytes(ot), xx, nbytes(xx), key, nbytes(key), sd,
        nbytes(sd), party, lf, nbytes(lf), rt, nbytes(rt), lt, nbytes(lt), n, m, es, N, None, 0)
    return oe, ot


def mic_keygen(be, alpha, s0, s1, left, right, n, m, es):
    N = len(alpha)
    I = len(left)
    ks = key_size(be, "dcfMICGetKeyDataSize", n, m, es, N)
    key = be.empty(ks)
    z = be.empty(N * I * es, DT[es])
    a, d0, d1, lf, rt = be.put(alpha), be.put(s0), be.put(s1), be.put(left), be.put(right)
    run(be, "dcfMICKeyGen", key, ks, z, nbytes(z), a, nbytes(a), d0, nbytes(d0), d1, nbytes(d1), lf, nbytes(lf), rt,
        nbytes(rt), n, m, es, N)
    return key, z


def mic_eval(be, x, key, z, seed, party, left, right, n, m, es):
    N = len(x)
    I = len(left)
    out = be.empty(N * es * I, DT[es])
    xx, zz, sd, lf, rt = be.put(x), be.put(z), be.put(seed), be.put(left), be.put(right)
    run(be, "dcfMICEval", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), zz, nbytes(zz), sd, nbytes(sd), party,
        lf, nbytes(lf), rt, nbytes(rt), n, m, es, N, None, 0)
    return out


def spline_keygen(be, alpha, s0, s1, coeff, degree, left, right, n, m, es):
    N = len(alpha)
    I = len(left)
    g = I * (degree + 1)
    ks = key_size(be, "dcfSplineGetKeyDataSize", degree, I, n, m, es, N)
    key = be.empty(ks)
    e = be.empty(N * g * es, DT[es])
    beta = be.empty(N * g * es, DT[es])
    a, d0, d1, cf, lf, rt = be.put(alpha), be.put(s0), be.put(s1), be.put(coeff), be.put(left), be.put(right)
    run(be, "dcfSplineKeyGen", key, ks, e, nbytes(e), beta, nbytes(beta), a, nbytes(a), d0, nbytes(d0), d1,
        nbytes(d1), cf, nbytes(cf), degree, lf, nbytes(lf), rt, nbytes(rt), I, n, m, es, N)
    return key, e, beta


def spline_eval(be, x, key, e, beta, seed, party, left, right, degree, n, m, es):
    N = len(x)
    I = len(left)
    cs = key_size(be, "dcfSplineGetCacheDataSize", degree, I, n, m, es, N)
    cache = be.empty(max(cs, 1))
    out = be.empty(N * es, DT[es])
    xx, sd, lf, rt = be.put(x), be.put(seed), be.put(left), be.put(right)
    e, beta = be.put(e), be.put(beta)
    run(be, "dcfSplineEval", out, nbytes(out), xx, nbytes(xx), key, nbytes(key), e, nbytes(e), beta, nbytes(beta),
        sd, nbytes(sd), party, lf, nbytes(lf), rt, nbytes(rt), I, degree, n, m, es, N, cache, cs)
    return out


def ottt_keygen(be, key_in, alpha, n, es):
    N = len(alpha)
    key = be.put(key_in.copy())
    a = be.put(alpha)
    run(be, "otttKeyGen", key, nbytes(key), a, nbytes(a), n, es, N)
    return key


def ottt_eval(be, x, key, party, lut, n, es):
    N = len(x)
    lutNum = len(lut) >> n
    oe = be.empty(N * es, DT[es])
    ot = be.empty(N * es * lutNum, DT[es])
    xx, lt = be.put(x), be.put(lut)
    run(be, "otttLutEval", oe, nbytes(oe), ot, nbytes(ot), xx, nbytes(xx), key, nbytes(key), party, lt, nbytes(lt), n,
        es, N)
    return oe, ot


def sub_key_rows(key, idx, sk, tail, N):
    """The key bytes of elements idx from a split key [first part: N x sk][tail: N x tail], in the same
    layout for len(idx) elements (DCF: sk = 16n, tail = (n g + 2 + g) es; DPF: sk = 16n, tail = (2 + g) es;
    Grotto: three parts, not covered)."""
    key = np.asarray(key)
    return np.concatenate([key[:N * sk].reshape(N, sk)[idx].ravel(), key[N * sk:N * (sk + tail)].reshape(N, tail)[idx].ravel()])


def spline_closed_form(X, alpha, left, right, coeff, d, n, m):
    """What two spline shares reconstruct to, from the mathematics alone (not the oracle): for the
    interval q holding x = (X - alpha) mod 2^n (unsigned, sorted disjoint intervals), f_q evaluated
    at the integer X - alpha, mod 2^m; 0 where no interval holds x. Evaluated in uint64 arithmetic,
    exact mod 2^m because 2^m divides 2^64."""
    X = np.asarray(X).astype(np.uint64)
    A = np.asarray(alpha).astype(np.uint64)
    x = (X - A) & np.uint64(mask(n))
    L = np.asarray(left).astype(np.uint64)
    R = np.asarray(right).astype(np.uint64)
    order = np.argsort(L)
    q = np.searchsorted(L[order], x, side="right") - 1
    hit = q >= 0
    qi = order[np.clip(q, 0, None)]
    hit &= x <= R[qi]
    C = np.asarray(coeff).astype(np.uint64).reshape(-1, d + 1)
    t = X - A  # the integer X - alpha, two's complement in 64 bits
    acc = np.zeros_like(X)
    pw = np.ones_like(X)
    with np.errstate(over="ignore"):
        for k in range(d + 1):
            acc += C[qi, k] * pw
            pw *= t
    return np.where(hit, acc & np.uint64(mask(m)), np.uint64(0))


def signed_nbit(v: int, bits: int) -> int:
    v &= mask(bits)
    return v - (1 << bits) if (bits and v >> (bits - 1)) else v


class DeviceBackend:
    """Runs the same argument lists through the device library's C ABI
    (FastFss_cuda_*) on HBM buffers allocated by torch on cuda:0."""

    name = "device"

    def __init__(self):
        import torch

        from fastfunctionsecretsharing_amd import _lib

        self.torch = torch
        self.lib = _lib.load()
        self.dev = torch.device("cuda:0")
        self._stream = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def put(self, a):
        if a is None:
            return None
        if hasattr(a, "data_ptr"):
            return a.contiguous()
        a = np.ascontiguousarray(a)
        if a.size == 0:
            return self.torch.empty(0, dtype=_TORCH_DT(self.torch)[a.dtype.itemsize], device=self.dev)
        t = self.torch.from_numpy(a.view(np.uint8).copy()).to(self.dev)
        return t.view(_TORCH_DT(self.torch)[a.dtype.itemsize]) if a.dtype != np.uint8 else t

    def empty(self, nbytes, dtype=np.uint8):
        it = np.dtype(dtype).itemsize
        return self.torch.zeros(nbytes // it, dtype=_TORCH_DT(self.torch)[it], device=self.dev)

    def get(self, t):
        self.torch.cuda.synchronize()
        it = t.element_size()
        return t.view(self.torch.uint8).cpu().numpy().view(DT[it])

    def call(self, name, *args):
        if name.startswith(("dcfGet", "dpfGet", "grottoGet", "dcfMICGet", "dcfSplineGet", "otttGet")):
            fn = getattr(self.lib, "FastFss_" + name)
            return fn(*[ptr(a) if (a is None or hasattr(a, "data_ptr")) else a for a in args])
        fn = getattr(self.lib, "FastFss_cuda_" + name)
        conv = [ptr(a) if (a is None or hasattr(a, "data_ptr")) else a for a in args]
        rc = fn(*conv, ctypes.addressof(self._stream))
        return rc


def _TORCH_DT(torch):
    # signed torch dtypes carry the same bit patterns and support every op needed
    return {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}


def to_np(be, a):
    return be.get(a) if hasattr(a, "data_ptr") else a
