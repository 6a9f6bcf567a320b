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


class _FakeCtx:
    """Records hfpx_tune_set calls; check() passes return codes through."""

    def __init__(self):
        self.set = {}
        self.handle = None
        ctx = self

        class Lib:
            @staticmethod
            def hfpx_tune_set(handle, knob, value):
                ctx.set[knob] = value
                return 0

        self.lib = Lib()

    def check(self, rc):
        assert rc == 0


def test_batch_row_apply_and_load(tmp_path):
    from hfpx3d_amd import _lib
    table = {"entries": [], "batch": {"knobs": {"BGEMM_F32": 2}, "mixed_min_n": 4096, "ms": 1.0, "tried": []}}
    ctx = _FakeCtx()
    assert autotune.apply_batch(table, ctx) == 4096
    assert ctx.set == {_lib.KNOB_BGEMM_F32: 2}
    ctx2 = _FakeCtx()
    assert autotune.apply_batch({"entries": []}, ctx2) is None and ctx2.set == {}
    p = str(tmp_path / "b.json")
    autotune.save({"batch": table["batch"]}, p)  # a table with only a batch row loads
    assert autotune.load(p)["batch"]["mixed_min_n"] == 4096


@pytest.mark.gpu
def test_batch_sweep_and_apply():
    """A batch sweep over 10 small cracks (720 to 2,340 DoF) with four
    candidates: the batch row is the fastest trial, the knob is left as it
    was, and a batch built with the row's settings solves every crack to a
    backward error below 1e-14."""
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, meshgen
    assert torch.cuda.is_available(), "GPU tests need a device"
    print("case: autotune batch sweep", flush=True)
    ctx = hx.get_context()
    cracks = []
    for c in meshgen.batch_cracks(10, 5, 40, 130):
        p = c["p"]
        cracks.append(dict(nods=c["nods"], conn=c["conn"], G=1.0 / (2.0 * (1.0 + c["nu"])), nu=c["nu"],
                           s_inf=[p, p, p, 0.0, 0.0, 0.0]))
    k = autotune.BATCH_KNOBS["BGEMM_F32"]
    before = hx.tune_get(k, ctx)
    cands = [{"BGEMM_F32": b, "mixed_min_n": m} for b in (2, 3) for m in (1024, 10 ** 6)]
    table = autotune.sweep_batch(cracks, candidates=cands, reps=1, ctx=ctx)
    row = table["batch"]
    assert row["ms"] == min(t["ms"] for t in row["tried"]) and math.isfinite(row["ms"])
    assert len(row["tried"]) == 4 and row["n_cracks"] == 10
    assert hx.tune_get(k, ctx) == before
    try:
        m = autotune.apply_batch(table, ctx)
        assert m == row["mixed_min_n"] and hx.tune_get(k, ctx) == row["knobs"]["BGEMM_F32"]
        bt = hx.Batch(cracks, mixed_min_n=m, ctx=ctx)
        bt.solve(n_refine=3, use_graph=True)
        for i in range(len(cracks)):
            x, berr, info = bt.solution(i)
            assert info == 0 and berr < 1e-14 and np.isfinite(x).all()
        bt.close()
    finally:
        hx.tune_set(k, before, ctx)
