"""Autotuner: table ranges and lookup on the CPU; a small sweep on the GPU,
whose table the LU then applies (no reference counterpart)."""
import math

import pytest

from hfpx3d_amd import autotune


def test_ranges_split_at_geometric_midpoints():
    r = autotune.ranges([16384, 4096, 65536])
    assert r[0] == (0, math.ceil(math.sqrt(4096 * 16384)))
    assert r[1] == (math.ceil(math.sqrt(4096 * 16384)), math.ceil(math.sqrt(16384 * 65536)))
    assert r[2] == (math.ceil(math.sqrt(16384 * 65536)), None)
    assert autotune.ranges([1000]) == [(0, None)]


def test_lookup_by_range_and_precision(tmp_path):
    table = {"entries": [
        {"precision": "f32", "n": 4096, "n_min": 0, "n_max": 8192, "knobs": {"LU_NB": 512}},
        {"precision": "f32", "n": 16384, "n_min": 8192, "n_max": None, "knobs": {"LU_NB": 2048, "GEMM_F32": 6}},
        {"precision": "f64", "n": 4096, "n_min": 0, "n_max": None, "knobs": {"LU_NB": 256}},
    ]}
    assert autotune.lookup(table, 100, "f32") == {"LU_NB": 512}
    assert autotune.lookup(table, 8191, "f32") == {"LU_NB": 512}
    assert autotune.lookup(table, 8192, "f32") == {"LU_NB": 2048, "GEMM_F32": 6}
    assert autotune.lookup(table, 10 ** 6, "f32") == {"LU_NB": 2048, "GEMM_F32": 6}
    assert autotune.lookup(table, 5, "f64") == {"LU_NB": 256}
    assert autotune.lookup({"entries": []}, 5, "f64") is None
    p = str(tmp_path / "t.json")
    autotune.save(table, p)
    assert autotune.load(p) == table
    with open(p, "w") as f:
        f.write("{}")
    with pytest.raises(ValueError):
        autotune.load(p)


@pytest.mark.gpu
def test_sweep_and_apply():
    """A small sweep (n = 600 and 2,000, two outer-block sizes per
    precision): one row per (size, precision), each the fastest of its
    trials; the knobs are left as they were; an LU then runs with its row's
    settings and still factorises correctly."""
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, _lib
    assert torch.cuda.is_available(), "GPU tests need a device"
    print("case: autotune sweep", flush=True)
    ctx = hx.get_context()
    before = {k: hx.tune_get(v, ctx) for k, v in autotune.KNOBS.items()}
    cands = {"f32": [{"LU_NB": 256}, {"LU_NB": 512, "GEMM_F32": 6}], "f64": [{"LU_NB": 256}, {"LU_NB": 512}]}
    table = autotune.sweep([600, 2000], ["f32", "f64"], candidates=cands, reps=1, ctx=ctx)
    assert len(table["entries"]) == 4
    for e in table["entries"]:
        assert e["ms"] == min(t["ms"] for t in e["tried"]) and math.isfinite(e["ms"])
    assert {k: hx.tune_get(v, ctx) for k, v in autotune.KNOBS.items()} == before
    try:
        autotune.use(table, ctx)
        n = 1500
        A = np.random.default_rng(1).normal(size=(n, n)) + np.sqrt(n) * np.eye(n)
        Ad, _ = hx.alloc_colmajor(n, n, torch.float32, ctx)
        Ad.copy_(torch.from_numpy(A))
        lu = hx.LU(Ad, ctx=ctx)
        want = autotune.lookup(table, n, "f32")
        for name, value in want.items():
            assert hx.tune_get(autotune.KNOBS[name], ctx) == value
        b = np.random.default_rng(2).normal(size=(n, 2))
        bd, _ = hx.alloc_colmajor(n, 2, torch.float32, ctx)
        bd.copy_(torch.from_numpy(b))
        x = lu.solve(bd).cpu().numpy().astype(np.float64)
        assert np.abs(A @ x - b).max() / (np.abs(A).max() * np.abs(x).max() * n) < 1e-6
    finally:
        ctx.autotune = None
        for k, v in before.items():
            hx.tune_set(autotune.KNOBS[k], v, ctx)
