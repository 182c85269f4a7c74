"""pyFastFss host-side behaviour without a GPU: the helper functions of the
reference package (lib/torch_extension/pyFastFss/__init__.py:5-70), size
queries equal to the oracle's, the reference's argument checks and reshaping,
and the routing of host tensors to the host-buffer entry points."""
import pytest
import torch

import fss_testlib as F
from fastfunctionsecretsharing_amd import _lib
from fastfunctionsecretsharing_amd import pyFastFss as P


def test_dtype_and_mod():
    assert [P.dtype(b) for b in (1, 8, 9, 16, 17, 32, 33, 64)] == [
        torch.int8, torch.int8, torch.int16, torch.int16, torch.int32, torch.int32, torch.int64, torch.int64]
    x = torch.tensor([-1, 5, 1 << 20, -(1 << 30)], dtype=torch.int32)
    assert P.mod(x, 16).tolist() == [0xFFFF, 5, 0, 0]
    assert torch.equal(P.mod(x, 32), x)
    out = torch.empty_like(x)
    assert P.mod(x, 8, out=out) is out and out.tolist() == [255, 5, 0, 0]
    w = torch.tensor([[-1, -1]], dtype=torch.int64)
    assert P.mod(w, 100).tolist() == [[-1, (1 << 36) - 1]]
    with pytest.raises(ValueError):
        P.mod(torch.zeros(3, dtype=torch.int64), 100)
    assert P.logical_shape(w, 100) == torch.Size([1]) and P.logical_numel(w, 100) == 1
    assert P.logical_shape(x, 32) == x.shape and P.logical_numel(x, 32) == 4


@pytest.mark.parametrize("args", [(32, 32, 1, 4, 100), (16, 8, 3, 2, 7), (64, 64, 1, 8, 1)])
def test_dcf_dpf_size_queries(args):
    for mine, ref in ((P.dcf_get_key_data_size, "dcfGetKeyDataSize"), (P.dpf_get_key_data_size, "dpfGetKeyDataSize")):
        assert mine(*args) == F.key_size(F.OracleBackend(), ref, *args)


def test_other_size_queries():
    O = F.OracleBackend()
    assert P.grotto_get_key_data_size(20, 4, 9) == F.key_size(O, "grottoGetKeyDataSize", 20, 4, 9)
    assert P.dcf_mic_get_key_data_size(16, 16, 2, 9) == F.key_size(O, "dcfMICGetKeyDataSize", 16, 16, 2, 9)
    assert P.spline_get_key_data_size(2, 4, 16, 16, 2, 9) == F.key_size(O, "dcfSplineGetKeyDataSize", 2, 4, 16, 16,
                                                                          2, 9)
    assert P.ottt_get_key_data_size(8, 16) == F.key_size(O, "otttGetKeyDataSize", 8, 16)
    with pytest.raises(RuntimeError, match="FastFss_dcfGetKeyDataSize failed"):
        P.dcf_get_key_data_size(65, 32, 1, 4, 1)


def test_config_round_trip():
    g = P.get_grid_dim()
    P.set_grid_dim(77)
    assert P.get_grid_dim() == 77
    P.set_grid_dim(g)
    t = P.get_fine_grain_parallel_grid_dim_threshold()
    P.set_fine_grain_parallel_grid_dim_threshold(5)
    assert P.get_fine_grain_parallel_grid_dim_threshold() == 5
    P.set_fine_grain_parallel_grid_dim_threshold(t)
    n = P.get_num_threads()
    P.set_num_threads(3)
    assert P.get_num_threads() == 3
    P.set_num_threads(n)


def _cpu_dcf_args(N=6, n=16, g=1):
    alpha = torch.zeros((2, N // 2), dtype=torch.int16)
    beta = torch.zeros(N * g, dtype=torch.int16)
    s = torch.zeros(16 * N, dtype=torch.uint8)
    return alpha, beta, s


class _RecordingLib:
    """Stands in for the device library: records each entry point called and returns an error code,
    so host tensors can be routed without touching a GPU."""

    def __init__(self, real, rc=-1):
        self.real, self.rc, self.calls = real, rc, []

    def __getattr__(self, name):
        if name.startswith(("FastFss_cpu_", "FastFss_cuda_")) and not name.endswith(("NumThreads", "GridDim")):
            def fn(*args):
                self.calls.append((name, len(args)))
                return None if name.endswith("prngInit") else self.rc
            return fn
        return getattr(self.real, name)


def test_cpu_tensors_route_to_host_buffer_api(monkeypatch):
    """CPU tensors go to FastFss_cpu_<op> (no stream argument), after the output is sized, as the
    reference binding dispatches on device (csrc/dcf/dcf.cpp:134-138); a failing code raises
    '<entry point> failed : <code>'."""
    rec = _RecordingLib(P._L())
    monkeypatch.setattr(P, "_L", lambda: rec)
    alpha, beta, s = _cpu_dcf_args()
    key = torch.empty(0, dtype=torch.uint8)
    with pytest.raises(RuntimeError, match="FastFss_cpu_dcfKeyGen failed : -1"):
        P.dcf_key_gen(key, alpha, beta, s, s, 16, 16, 1)
    # the key is sized before the entry point is called, as in the reference
    assert key.numel() == P.dcf_get_key_data_size(16, 16, 1, 2, 6)
    out = torch.empty(0, dtype=torch.int16)
    with pytest.raises(RuntimeError, match="FastFss_cpu_dcfEval failed"):
        P.dcf_eval(out, alpha, key, s, 0, 16, 16, 2)
    assert tuple(out.shape) == (2, 3, 2)
    out = torch.empty(0, dtype=torch.int16)
    with pytest.raises(RuntimeError, match="FastFss_cpu_dpfEvalAll failed"):
        P.dpf_eval_all(out, alpha, key, s, 0, 4, 16, 1)
    assert tuple(out.shape) == (2, 3, 16)
    with pytest.raises(RuntimeError, match="FastFss_cpu_prngInit failed"):
        P.Prng()
    # the host-buffer signatures: the device ones without the trailing stream pointer
    cuda_args = {n: len(k) for n, (k, _) in _lib.SIGNATURES.items()}
    for name, nargs in rec.calls:
        assert name.startswith("FastFss_cpu_")
        if not name.endswith("prngInit"):
            assert nargs == cuda_args[name.replace("cpu", "cuda")] - 1, name


def test_argument_checks():
    alpha, beta, s = _cpu_dcf_args()
    key = torch.empty(0, dtype=torch.uint8)
    with pytest.raises(RuntimeError):
        P.dcf_key_gen(key, alpha, beta.to(torch.int32), s, s, 16, 16, 1)  # beta dtype differs from alpha
    with pytest.raises(RuntimeError):
        P.dcf_key_gen(key.to(torch.int8), alpha, beta, s, s, 16, 16, 1)  # key must be uint8
    with pytest.raises(RuntimeError):
        P.dcf_eval(torch.empty(0, dtype=torch.int16), alpha.t(), key, s, 0, 16, 16, 1)  # not contiguous
    with pytest.raises(RuntimeError):
        P.dcf_eval(torch.empty(0, dtype=torch.int32), alpha, key, s, 0, 16, 16, 1)  # out dtype differs
    with pytest.raises(RuntimeError):
        P.grotto_eval(torch.empty(0, dtype=torch.int16), alpha, key, s.to(torch.int8), False, 0, 16)


def test_top_level_import_is_the_same_module():
    """Callers of the reference keep `import pyFastFss` (lib/torch_extension/setup.py:28-96)."""
    import pyFastFss

    assert pyFastFss is P
    assert pyFastFss.dcf_eval is P.dcf_eval and pyFastFss.Prng is P.Prng
