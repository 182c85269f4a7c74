"""Loader for the device library ``libfastfss_hip.so`` and its C ABI.

The library is the product: every ``FastFss_cuda_*`` entry point runs HIP
kernels for gfx950.  There is no fallback path; if the library is missing the
import fails with an error naming the build command.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfastfss_hip.so"

# argument kinds: p = pointer, s = size_t, i = int, b = bool
_EVAL = "pspspsps" + "i" + "sssss" + "ps"
SIGNATURES = {
    # name: (argument kinds, restype)
    "FastFss_cuda_dcfKeyGen": ("pspspspsps" + "sssss" + "p", "int"),
    "FastFss_cuda_dcfEval": (_EVAL + "p", "int"),
    "FastFss_cuda_dpfKeyGen": ("pspspspsps" + "sssss" + "p", "int"),
    "FastFss_cuda_dpfEval": (_EVAL + "p", "int"),
    "FastFss_cuda_dpfEvalAll": (_EVAL + "p", "int"),
    "FastFss_cuda_dpfEvalMulti": ("pspspsps" + "i" + "ps" + "sssss" + "ps" + "p", "int"),
    "FastFss_cuda_grottoKeyGen": ("pspspsps" + "sss" + "p", "int"),
    "FastFss_cuda_grottoEqEval": ("pspspsps" + "i" + "sss" + "ps" + "p", "int"),
    "FastFss_cuda_grottoEval": ("pspspsps" + "b" + "i" + "sss" + "ps" + "p", "int"),
    "FastFss_cuda_grottoMICEval": ("pspspsps" + "i" + "psps" + "sss" + "ps" + "p", "int"),
    "FastFss_cuda_grottoIntervalLutEval": ("pspspspsps" + "i" + "pspsps" + "ssss" + "ps" + "p", "int"),
    "FastFss_cuda_dcfMICKeyGen": ("pspspspspspsps" + "ssss" + "p", "int"),
    "FastFss_cuda_dcfMICEval": ("pspspspsps" + "i" + "psps" + "ssss" + "ps" + "p", "int"),
    "FastFss_cuda_dcfSplineKeyGen": ("pspspspspspsps" + "s" + "psps" + "sssss" + "p", "int"),
    "FastFss_cuda_dcfSplineEval": ("pspspspspsps" + "i" + "psps" + "ssssss" + "ps" + "p", "int"),
    "FastFss_cuda_otttKeyGen": ("psps" + "sss" + "p", "int"),
    "FastFss_cuda_otttLutEval": ("pspspsps" + "i" + "ps" + "sss" + "p", "int"),
    "FastFss_cuda_prngInit": ("", "ptr"),
    "FastFss_cuda_prngRelease": ("p", None),
    "FastFss_cuda_prngSetCurrentSeed": ("ppp", "int"),
    "FastFss_cuda_prngGetCurrentSeed": ("ppp", "int"),
    "FastFss_cuda_prngGen": ("ppsss" + "p", "int"),
    "FastFss_cuda_setGridDim": ("i", "int"),
    "FastFss_cuda_getGridDim": ("", "int"),
    "FastFss_cuda_setFineGrainParallelGridDimThreshold": ("i", "int"),
    "FastFss_cuda_getFineGrainParallelGridDimThreshold": ("", "int"),
    "FastFss_dcfKeyZip": ("psps" + "sssss", "int"),
    "FastFss_dcfKeyUnzip": ("psps" + "sssss", "int"),
    "FastFss_dcfGetKeyDataSize": ("psssss", "int"),
    "FastFss_dcfGetZippedKeyDataSize": ("psssss", "int"),
    "FastFss_dcfGetCacheDataSize": ("psssss", "int"),
    "FastFss_dpfKeyZip": ("psps" + "sssss", "int"),
    "FastFss_dpfKeyUnzip": ("psps" + "sssss", "int"),
    "FastFss_dpfGetKeyDataSize": ("psssss", "int"),
    "FastFss_dpfGetZippedKeyDataSize": ("psssss", "int"),
    "FastFss_dpfGetCacheDataSize": ("psss", "int"),
    "FastFss_grottoKeyZip": ("psps" + "sss", "int"),
    "FastFss_grottoKeyUnzip": ("psps" + "sss", "int"),
    "FastFss_grottoGetKeyDataSize": ("psss", "int"),
    "FastFss_grottoGetZippedKeyDataSize": ("psss", "int"),
    "FastFss_grottoGetCacheDataSize": ("psss", "int"),
    "FastFss_dcfMICKeyZip": ("psps" + "ssss", "int"),
    "FastFss_dcfMICKeyUnzip": ("psps" + "ssss", "int"),
    "FastFss_dcfMICGetCacheDataSize": ("pssss", "int"),
    "FastFss_dcfMICGetKeyDataSize": ("pssss", "int"),
    "FastFss_dcfMICGetZippedKeyDataSize": ("pssss", "int"),
    "FastFss_dcfSplineKeyZip": ("psps" + "ssssss", "int"),
    "FastFss_dcfSplineKeyUnzip": ("psps" + "ssssss", "int"),
    "FastFss_dcfSplineGetCacheDataSize": ("pssssss", "int"),
    "FastFss_dcfSplineGetKeyDataSize": ("pssssss", "int"),
    "FastFss_dcfSplineGetZippedKeyDataSize": ("pssssss", "int"),
    "FastFss_otttGetKeyDataSize": ("pss", "int"),
    # the host-buffer API (include/FastFss/cpu/*.h): the device entry points' arguments without the stream
    "FastFss_cpu_setNumThreads": ("i", "int"),
    "FastFss_cpu_getNumThreads": ("", "int"),
    "FastFss_cpu_prngInit": ("", "ptr"),
    "FastFss_cpu_prngRelease": ("p", None),
    "FastFss_cpu_prngSetCurrentSeed": ("ppp", "int"),
    "FastFss_cpu_prngGetCurrentSeed": ("ppp", "int"),
    "FastFss_cpu_prngGen": ("ppsss", "int"),
    # MI355X-only entry points (include/FastFss/hip/*.h)
    "FastFss_hip_batchPlanCreate": ("pps", "int"),
    "FastFss_hip_batchPlanLaunch": ("pp", "int"),
    "FastFss_hip_batchPlanCapture": ("pp", "int"),
    "FastFss_hip_batchPlanReplay": ("pp", "int"),
    "FastFss_hip_batchPlanTileCount": ("p", "size"),
    "FastFss_hip_batchPlanDestroy": ("p", None),
    "FastFss_hip_batchEval": ("psp", "int"),
    "FastFss_hip_batchSetDefaultTile": ("isu", "int"),
    "FastFss_hip_batchGetDefaultTile": ("is", "uint"),
    "FastFss_hip_batchSetBlocksPerCu": ("i", "int"),
    "FastFss_hip_setEvalVariant": ("i", "int"),
    "FastFss_hip_getEvalVariantCount": ("", "int"),
    "FastFss_hip_setEvalVariantForElementSize": ("si", "int"),
    "FastFss_hip_dcfEvalNoAlloc": (_EVAL + "p", "int"),
    "FastFss_hip_grottoEvalFullDomain": ("pspsps" + "iii" + "sss" + "p", "int"),
    "FastFss_hip_grottoFullSetLaneTarget": ("s", "int"),
    "FastFss_hip_serverCreate": ("psiii", "int"),
    "FastFss_hip_serverSubmit": ("ppp", "int"),
    "FastFss_hip_serverQuery": ("pq", "int"),
    "FastFss_hip_serverWait": ("pqp", "int"),
    "FastFss_hip_serverRelease": ("pq", "int"),
    "FastFss_hip_serverDrain": ("p", "int"),
    "FastFss_hip_serverStats": ("pp", "int"),
    "FastFss_hip_serverDestroy": ("p", None),
    "FastFss_hip_keyStoreCreate": ("ps", "int"),
    "FastFss_hip_keyStoreDestroy": ("p", "int"),
    "FastFss_hip_keyStorePut": ("ppspp", "int"),
    "FastFss_hip_keyStoreReserve": ("pspp", "int"),
    "FastFss_hip_keyStoreGet": ("pqpp", "int"),
    "FastFss_hip_keyStoreRelease": ("pqp", "int"),
    "FastFss_hip_keyStoreStats": ("ppppp", "int"),
}


class ServeRequest(ctypes.Structure):
    """ctypes mirror of FastFssServeRequest (include/FastFss/hip/serve.h)."""

    _fields_ = [
        ("op", ctypes.c_int),
        ("partyId", ctypes.c_int),
        ("bitWidthIn", ctypes.c_size_t),
        ("bitWidthOut", ctypes.c_size_t),
        ("elementSize", ctypes.c_size_t),
        ("elementNum", ctypes.c_size_t),
        ("sharedOut", ctypes.c_void_p),
        ("maskedX", ctypes.c_void_p),
        ("key", ctypes.c_void_p),
        ("keyDataSize", ctypes.c_size_t),
        ("seed", ctypes.c_void_p),
        ("sharedZ", ctypes.c_void_p),
        ("sharedBeta", ctypes.c_void_p),
        ("leftEndpoints", ctypes.c_void_p),
        ("rightEndpoints", ctypes.c_void_p),
        ("intervalNum", ctypes.c_size_t),
        ("degree", ctypes.c_size_t),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("maxInflightSeen", ctypes.c_uint64),
        ("elementsDone", ctypes.c_uint64),
        ("busySeconds", ctypes.c_double),
        ("launches", ctypes.c_uint64),
        ("fusedRequests", ctypes.c_uint64),
    ]


SERVE_DCF, SERVE_MIC, SERVE_SPLINE, SERVE_DPF = 0, 1, 2, 3


class BatchGroup(ctypes.Structure):
    """ctypes mirror of FastFssBatchGroup (include/FastFss/hip/batch.h)."""

    _fields_ = [
        ("op", ctypes.c_int),
        ("partyId", ctypes.c_int),
        ("bitWidthIn", ctypes.c_size_t),
        ("bitWidthOut", ctypes.c_size_t),
        ("elementSize", ctypes.c_size_t),
        ("elementNum", ctypes.c_size_t),
        ("sharedOut", ctypes.c_void_p),
        ("maskedX", ctypes.c_void_p),
        ("key", ctypes.c_void_p),
        ("seed", ctypes.c_void_p),
        ("keyDataSize", ctypes.c_size_t),
        ("tileElems", ctypes.c_uint32),
    ]


GROTTO_FULL_LT, GROTTO_FULL_LE, GROTTO_FULL_EQ = 0, 1, 2

BATCH_DCF_EVAL = 0
BATCH_DPF_EVAL = 1

for _name in [n for n in SIGNATURES if n.startswith("FastFss_cuda_") and "prng" not in n and "Grid" not in n]:
    _kinds, _res = SIGNATURES[_name]
    SIGNATURES[_name.replace("FastFss_cuda_", "FastFss_cpu_")] = (_kinds[:-1], _res)

_KIND = {"p": ctypes.c_void_p, "s": ctypes.c_size_t, "i": ctypes.c_int, "b": ctypes.c_bool, "u": ctypes.c_uint32,
         "q": ctypes.c_uint64}
_RES = {"int": ctypes.c_int, "ptr": ctypes.c_void_p, "size": ctypes.c_size_t, "uint": ctypes.c_uint32, None: None}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path: Path | str | None = None):
    """Load (once) and return the ctypes handle with argument types bound."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    # torch bundles its own HIP runtime under the same sonames as the system's; if this library loaded
    # first, its libamdhip64.so.7 would be the one a later `import torch` finds, and torch would then
    # report no GPU.  Importing torch first (when installed) makes both use torch's runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not p.exists():
        raise LibraryMissing(
            f"{p} is not built; run `python -m fastfunctionsecretsharing_amd.build` "
            "(the MI355X kernels are the only implementation)")
    lib = ctypes.CDLL(str(p))
    for name, (kinds, res) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = [_KIND[k] for k in kinds]
        fn.restype = _RES[res]
    if path is None:
        _lib = lib
    return lib


def exported_symbols(path: Path | str | None = None):
    lib = load(path)
    return {name for name in SIGNATURES if getattr(lib, name, None) is not None}
