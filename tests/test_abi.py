"""The device library builds for gfx950, loads, and exports every entry point
declared under include/FastFss (no compute calls: this runs without a GPU)."""
import re
from pathlib import Path

import pytest

from fastfunctionsecretsharing_amd import _lib

REPO = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (REPO / "include" / "FastFss").rglob("*.h"):
        for m in re.finditer(r"FAST_FSS_API\s+[\w\s\*]+?\b(FastFss_\w+)\s*\(", h.read_text()):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libfastfss_hip.so is not built (python -m fastfunctionsecretsharing_amd.build)")
    return _lib.load()


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in sorted(declared_symbols()) if getattr(lib, n, None) is None]
    assert not missing, missing


def test_signature_table_covers_headers():
    assert declared_symbols() <= set(_lib.SIGNATURES), sorted(declared_symbols() - set(_lib.SIGNATURES))


def test_size_queries_match_oracle(lib):
    import ctypes

    from fss_testlib import ocall

    cases = [
        ("dcfGetKeyDataSize", (32, 32, 1, 4, 1 << 24)),
        ("dcfGetKeyDataSize", (16, 16, 1, 2, 1 << 16)),
        ("dcfGetKeyDataSize", (33, 32, 1, 4, 5)),
        ("dcfGetKeyDataSize", (8, 8, 1, 3, 5)),
        ("dcfGetCacheDataSize", (24, 32, 48, 4, 7)),
        ("dpfGetKeyDataSize", (32, 32, 1, 4, 1 << 24)),
        ("dpfGetKeyDataSize", (0, 0, 1, 4, 3)),
        ("dpfGetCacheDataSize", (20, 4, 9)),
        ("grottoGetKeyDataSize", (16, 2, 4096)),
        ("grottoGetKeyDataSize", (5, 2, 4096)),
        ("grottoGetKeyDataSize", (16, 3, 4096)),
        ("grottoGetCacheDataSize", (16, 2, 4096)),
        ("dcfMICGetKeyDataSize", (24, 32, 4, 1 << 22)),
        ("dcfMICGetCacheDataSize", (24, 32, 4, 11)),
        ("dcfSplineGetKeyDataSize", (2, 16, 24, 32, 4, 1 << 22)),
        ("dcfSplineGetCacheDataSize", (2, 16, 24, 32, 4, 1 << 22)),
        ("otttGetKeyDataSize", (10, 77)),
        ("otttGetKeyDataSize", (2, 77)),
    ]
    for name, args in cases:
        a, b = ctypes.c_size_t(0), ctypes.c_size_t(0)
        ra = getattr(lib, "FastFss_" + name)(ctypes.addressof(a), *args)
        rb = ocall(name, ctypes.addressof(b), *args)
        assert (ra, a.value) == (rb, b.value), (name, args)


def test_zip_stubs_report_runtime_error(lib):
    assert lib.FastFss_dcfKeyZip(None, 0, None, 0, 8, 8, 1, 1, 1) == -1
    assert lib.FastFss_dpfKeyUnzip(None, 0, None, 0, 8, 8, 1, 1, 1) == -1
    assert lib.FastFss_grottoKeyZip(None, 0, None, 0, 8, 1, 1) == -1


def test_config_roundtrip(lib):
    assert lib.FastFss_cuda_setGridDim(0) == -2
    old = lib.FastFss_cuda_getGridDim()
    assert lib.FastFss_cuda_setGridDim(77) == 0 and lib.FastFss_cuda_getGridDim() == 77
    lib.FastFss_cuda_setGridDim(old)
