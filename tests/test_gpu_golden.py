"""The device library reproduces every committed golden vector (keys, shares,
PRNG stream) bit for bit."""
import ctypes

import numpy as np
import pytest

from fss_testlib import DeviceBackend
from golden_cases import REPLAYS, load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    return DeviceBackend()


@pytest.mark.parametrize("name", sorted(REPLAYS))
def test_device_reproduces_golden(dev, name):
    REPLAYS[name](dev)


def test_device_prng_golden(dev):
    z = load("prng")
    lib = dev.lib
    g = lib.FastFss_cuda_prngInit()
    assert lib.FastFss_cuda_prngSetCurrentSeed(g, z["seed"].tobytes(), z["counter"].tobytes()) == 0
    out = dev.empty(z["out_u32"].nbytes, np.uint32)
    assert lib.FastFss_cuda_prngGen(g, out.data_ptr(), 32, 4, out.numel(), ctypes.addressof(dev._stream)) == 0
    assert np.array_equal(dev.get(out), z["out_u32"])
    lib.FastFss_cuda_prngRelease(g)
