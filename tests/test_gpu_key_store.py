"""Device-resident key store (include/FastFss/hip/key_store.h): keys put from
host and device memory, or generated straight into a reserved slot, evaluate
through every entry point exactly as the caller's buffer does; released slots
are reused only after the release stream's work, and free space merges back
into one extent."""
import ctypes
import threading

import numpy as np
import pytest

from fss_testlib import (DT, DeviceBackend, OracleBackend, dcf_eval, dcf_keygen, dpf_eval, dpf_keygen, key_size,
                         rand_elems, rand_seeds, run)

pytestmark = pytest.mark.gpu


class DevPtr:
    """A device buffer by address and size, for the test library's argument lists."""

    def __init__(self, addr, size):
        self.addr, self.size = addr, size

    def data_ptr(self):
        return self.addr

    def numel(self):
        return self.size

    def element_size(self):
        return 1

    def contiguous(self):
        return self


@pytest.fixture(scope="module")
def dev():
    return DeviceBackend()


class Store:
    def __init__(self, lib, capacity):
        self.lib = lib
        self.h = ctypes.c_void_p()
        assert lib.FastFss_hip_keyStoreCreate(ctypes.addressof(self.h), capacity) == 0

    def put(self, addr, size, sp=None):
        h = ctypes.c_uint64(0)
        rc = self.lib.FastFss_hip_keyStorePut(self.h, addr, size, ctypes.addressof(h), sp)
        return rc, h.value

    def reserve(self, size):
        h, p = ctypes.c_uint64(0), ctypes.c_void_p()
        rc = self.lib.FastFss_hip_keyStoreReserve(self.h, size, ctypes.addressof(h), ctypes.addressof(p))
        return rc, h.value, p.value

    def get(self, handle):
        p, s = ctypes.c_void_p(), ctypes.c_size_t()
        rc = self.lib.FastFss_hip_keyStoreGet(self.h, handle, ctypes.addressof(p), ctypes.addressof(s))
        return rc, p.value, s.value

    def release(self, handle, sp=None):
        return self.lib.FastFss_hip_keyStoreRelease(self.h, handle, sp)

    def stats(self):
        v = [ctypes.c_size_t() for _ in range(4)]
        assert self.lib.FastFss_hip_keyStoreStats(self.h, *[ctypes.addressof(x) for x in v]) == 0
        return [x.value for x in v]

    def destroy(self):
        assert self.lib.FastFss_hip_keyStoreDestroy(self.h) == 0


def test_put_get_eval_equals_caller_buffer(dev):
    """DCF and DPF keys put from device and from host memory evaluate to the
    same shares as the original buffer, and to the oracle's."""
    torch, lib = dev.torch, dev.lib
    rng = np.random.default_rng(41)
    N, n, m, g, es = 3000, 32, 32, 1, 4
    alpha, x = rand_elems(rng, N, n, es), rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    st = Store(lib, 64 << 20)
    try:
        for keygen, evalf in ((dcf_keygen, dcf_eval), (dpf_keygen, dpf_eval)):
            kd = keygen(dev, alpha, None, s0, s1, n, m, g, es)
            torch.cuda.synchronize()
            host = dev.get(kd).copy()
            rc, hd = st.put(kd.data_ptr(), kd.numel())
            assert rc == 0 and hd != 0
            rc, hh = st.put(host.ctypes.data, host.nbytes)
            assert rc == 0 and hh not in (0, hd)
            want = dev.get(evalf(dev, x, kd, s1, 1, n, m, g, es))
            # an element's shares depend on its own key rows only: the oracle's key for the first 256
            # elements gives the first 256 shares
            orc_key = keygen(OracleBackend(), alpha[:256], None, s0[:256 * 16], s1[:256 * 16], n, m, g, es)
            assert np.array_equal(evalf(OracleBackend(), x[:256], orc_key, s1[:256 * 16], 1, n, m, g, es),
                                  want[:256])
            for h in (hd, hh):
                rc, p, size = st.get(h)
                assert rc == 0 and size == host.nbytes and p % 256 == 0
                y = dev.get(evalf(dev, x, DevPtr(p, size), s1, 1, n, m, g, es))
                assert np.array_equal(y, want)
            assert st.release(hd) == 0 and st.release(hh) == 0
        torch.cuda.synchronize()
        used, free, count, largest = st.stats()
        assert used == 0 and count == 0 and free == 64 << 20 and largest == 64 << 20
    finally:
        st.destroy()


def test_reserve_and_keygen_into_the_store(dev):
    """A keyGen writing straight into a reserved slot makes the oracle's key: both parties'
    shares equal the oracle's on every element."""
    torch, lib = dev.torch, dev.lib
    rng = np.random.default_rng(42)
    N, n, m, g, es = 1000, 24, 16, 1, 4
    alpha = rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    ks = key_size(dev, "dcfGetKeyDataSize", n, m, g, es, N)
    st = Store(lib, 8 << 20)
    try:
        rc, h, p = st.reserve(ks)
        assert rc == 0 and p % 256 == 0
        key = DevPtr(p, ks)
        a, d0, d1 = dev.put(alpha), dev.put(s0), dev.put(s1)
        run(dev, "dcfKeyGen", key, ks, a, a.numel() * es, None, 0, d0, N * 16, d1, N * 16, n, m, g, es, N)
        torch.cuda.synchronize()
        rc2, p2, size2 = st.get(h)
        assert rc2 == 0 and p2 == p and size2 == ks
        want_key = dcf_keygen(OracleBackend(), alpha, None, s0, s1, n, m, g, es)
        x = rand_elems(rng, N, n, es)
        for party, seed in ((0, s0), (1, s1)):
            y = dev.get(dcf_eval(dev, x, key, seed, party, n, m, g, es))
            yo = dcf_eval(OracleBackend(), x, want_key, seed, party, n, m, g, es)
            assert np.array_equal(y, yo), party
        assert st.release(h) == 0
    finally:
        st.destroy()


def test_release_is_stream_ordered_and_space_merges(dev):
    torch, lib = dev.torch, dev.lib
    st = Store(lib, 1 << 20)
    stream = torch.cuda.Stream()
    sp_holder = ctypes.c_void_p(stream.cuda_stream)
    sp = ctypes.addressof(sp_holder)
    try:
        src = torch.arange(300 * 1024, dtype=torch.uint8, device="cuda:0")
        rc, a = st.put(src.data_ptr(), 300 * 1024)
        assert rc == 0
        rc, b = st.put(src.data_ptr(), 300 * 1024)
        assert rc == 0
        _, pa, _ = st.get(a)
        _, pb, _ = st.get(b)
        assert pb == pa + 300 * 1024  # first fit, 256-byte aligned
        # a third slot of that size does not fit beside the other two (1 MiB)
        rc, c = st.put(src.data_ptr(), 500 * 1024)
        assert rc != 0 and c == 0
        # release a on a stream that is still busy: the slot comes back once the stream's work is done
        with torch.cuda.stream(stream):
            torch.cuda._sleep(20_000_000)
        assert st.release(a, sp) == 0
        used, free, count, largest = st.stats()
        assert count == 1 and used == 300 * 1024
        stream.synchronize()
        rc, d = st.put(src.data_ptr(), 200 * 1024)
        assert rc == 0
        _, pd, _ = st.get(d)
        assert pd == pa  # reused from the released slot, first fit
        assert st.get(a)[0] != 0  # released handles are gone
        assert st.release(a) != 0
        assert st.release(b) == 0 and st.release(d) == 0
        torch.cuda.synchronize()
        used, free, count, largest = st.stats()
        assert (used, count) == (0, 0) and free == largest == 1 << 20
        # a put that needs a pending release waits for it instead of failing
        rc, e = st.put(src.data_ptr(), 900 * 1024)
        assert rc == 0
        with torch.cuda.stream(stream):
            torch.cuda._sleep(20_000_000)
        assert st.release(e, sp) == 0
        rc, f = st.put(src.data_ptr(), 900 * 1024)
        assert rc == 0
        _, pf, _ = st.get(f)
        assert pf == pa
        torch.cuda.synchronize()
        assert st.release(f) == 0
    finally:
        torch.cuda.synchronize()
        st.destroy()


def test_invalid_calls(dev):
    lib = dev.lib
    h = ctypes.c_void_p()
    assert lib.FastFss_hip_keyStoreCreate(ctypes.addressof(h), 0) != 0
    assert lib.FastFss_hip_keyStoreCreate(None, 1024) != 0
    st = Store(lib, 4096)
    try:
        assert st.put(None, 16)[0] != 0
        buf = np.zeros(16, np.uint8)
        assert st.put(buf.ctypes.data, 0)[0] != 0
        assert st.put(buf.ctypes.data, 8192)[0] != 0  # larger than the store
        # sizes whose round-up to 256 bytes would wrap around: rejected before any copy
        assert st.put(buf.ctypes.data, 2**64 - 1)[0] != 0
        assert st.put(buf.ctypes.data, 2**64 - 200)[0] != 0
        assert st.reserve(2**64 - 1)[0] != 0
        assert st.get(12345)[0] != 0
        assert st.release(12345) != 0
        assert st.reserve(0)[0] != 0
        assert st.stats()[0] == 0 and st.stats()[2] == 0  # nothing was handed out
    finally:
        st.destroy()
    assert lib.FastFss_hip_keyStoreCreate(ctypes.addressof(h), 2**64 - 1) != 0
    assert lib.FastFss_hip_keyStoreCreate(ctypes.addressof(h), 2**64 - 100) != 0


def test_concurrent_threads(dev):
    """Host threads put, read back through an eval-free copy check, and release at once."""
    torch, lib = dev.torch, dev.lib
    st = Store(lib, 16 << 20)
    errors = []
    src = [torch.full((64 * 1024 + 17 * t,), t + 1, dtype=torch.uint8, device="cuda:0") for t in range(4)]
    torch.cuda.synchronize()

    def worker(t):
        try:
            for it in range(50):
                rc, h = st.put(src[t].data_ptr(), src[t].numel())
                if rc:
                    errors.append(("put", t, it, rc))
                    return
                rc, p, size = st.get(h)
                if rc or size != src[t].numel():
                    errors.append(("get", t, it, rc))
                    return
                if st.release(h):
                    errors.append(("release", t, it))
                    return
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    assert not errors, errors
    used, free, count, largest = st.stats()
    assert (used, count) == (0, 0) and largest == 16 << 20
    st.destroy()


def test_served_request_with_a_stored_key(dev):
    """A request whose key lives in the store serves to the per-call output."""
    torch, lib = dev.torch, dev.lib
    from fastfunctionsecretsharing_amd import _lib

    rng = np.random.default_rng(44)
    N, n, m, g, es = 5000, 16, 16, 1, 2
    alpha, x = rand_elems(rng, N, n, es), rand_elems(rng, N, n, es)
    s0, s1 = rand_seeds(rng, N), rand_seeds(rng, N)
    kd = dcf_keygen(dev, alpha, None, s0, s1, n, m, g, es)
    torch.cuda.synchronize()
    st = Store(lib, 4 << 20)
    srv = ctypes.c_void_p()
    assert lib.FastFss_hip_serverCreate(ctypes.addressof(srv), 1 << 30, 4, 2, 0) == 0
    try:
        rc, h = st.put(kd.data_ptr(), kd.numel())
        assert rc == 0
        torch.cuda.synchronize()
        _, p, size = st.get(h)
        want = dev.get(dcf_eval(dev, x, kd, s0, 0, n, m, g, es))
        xd, sd = dev.put(x), dev.put(s0)
        out = torch.empty(N, dtype=torch.int16, device="cuda:0")
        req = _lib.ServeRequest(op=_lib.SERVE_DCF, partyId=0, bitWidthIn=n, bitWidthOut=m, elementSize=es,
                                elementNum=N, sharedOut=out.data_ptr(), maskedX=xd.data_ptr(), key=p,
                                keyDataSize=size, seed=sd.data_ptr())
        ticket = ctypes.c_uint64()
        assert lib.FastFss_hip_serverSubmit(srv, ctypes.byref(req), ctypes.addressof(ticket)) == 0
        lat = ctypes.c_double()
        assert lib.FastFss_hip_serverWait(srv, ticket.value, ctypes.addressof(lat)) == 0
        assert np.array_equal(out.view(torch.uint8).cpu().numpy().view(DT[es]), want)
        assert st.release(h) == 0
    finally:
        lib.FastFss_hip_serverDestroy(srv)
        torch.cuda.synchronize()
        st.destroy()
