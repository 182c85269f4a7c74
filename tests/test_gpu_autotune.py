"""The autotuner picks valid launch shapes on the device, applies them, and the
tuned shapes leave every output unchanged (tiles and eval variants only move
work between lanes)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fastfunctionsecretsharing_amd import _lib, autotune, workloads as W

    torch.cuda.set_device(0)
    lib = _lib.load()
    saved = {(op, es): lib.FastFss_hip_batchGetDefaultTile(op, es) for op in (0, 1) for es in (1, 2, 4, 8)}
    yield torch, lib, autotune, W
    for (op, es), t in saved.items():
        lib.FastFss_hip_batchSetDefaultTile(op, es, t)
    for es in (1, 2, 4, 8):
        lib.FastFss_hip_setEvalVariantForElementSize(es, -1)


def test_tile_tuning_keeps_outputs(env, tmp_path):
    torch, lib, autotune, W = env
    groups = W.ragged_batch(groups=24, seed=3, lo=6, hi=12)
    holder, sp = W.stream_ptr()
    ref = W.alloc_outputs(groups)
    W.eval_plain(groups, 0, ref, sp)
    cands = (64, 256, 1024)
    table, ms, secs = autotune.tune_batch_tiles(groups, candidates=cands, reps=1)
    assert ms > 0 and secs > 0
    assert set(table) == {(g.op, g.es) for g in groups}
    for (op, es), t in table.items():
        assert t > 0 and t % 64 == 0  # a candidate, or the default it did not beat
        assert lib.FastFss_hip_batchGetDefaultTile(op, es) == t
    outs = W.alloc_outputs(groups)
    desc = W.batch_descriptors(groups, 0, outs)
    assert lib.FastFss_hip_batchEval(desc, len(groups), sp) == 0
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(outs, ref))
    path = tmp_path / "tune.json"
    autotune.save(path, tiles=table)
    for (op, es) in table:
        lib.FastFss_hip_batchSetDefaultTile(op, es, 64)
    entry = autotune.load_and_apply(path)
    assert len(entry["tiles"]) == len(table)
    for (op, es), t in table.items():
        assert lib.FastFss_hip_batchGetDefaultTile(op, es) == t
    del holder


def test_variant_tuning_keeps_outputs(env):
    torch, lib, autotune, W = env
    dev = torch.device("cuda:0")
    holder, sp = W.stream_ptr()
    gen = torch.Generator(device=dev).manual_seed(5)
    N, n = 1 << 14, 32
    g = W.make_group(0, n, N, gen, dev, sp)

    def run():
        out = torch.empty(N, dtype=torch.int32, device=dev)
        rc = lib.FastFss_cuda_dcfEval(out.data_ptr(), 4 * N, g.masked.data_ptr(), 4 * N, g.key.data_ptr(), g.key_size,
                                      g.seeds[0].data_ptr(), 16 * N, 0, n, n, 1, 4, N, None, 0, sp)
        assert rc == 0
        torch.cuda.synchronize()
        return out

    before = run()
    best, ms = autotune.tune_eval_variants(widths=(16, 32), log2n=14, reps=1)
    assert set(best) == {2, 4}
    assert all(0 <= v < lib.FastFss_hip_getEvalVariantCount() for v in best.values())
    assert torch.equal(run(), before)
    del holder
