"""hfpx_stress_noalloc, the allocation-free form of hfpx_stress, through the
C ABI: two calls in a row (the second on a smaller mesh and fewer points)
against hfpx_stress on the same inputs (bitwise: the same kernels with the
same element chunks) and against the oracle (1e-12 of the largest stress),
and its refusal of a problem beyond its fixed workspace."""
import ctypes

import numpy as np
import pytest

import oracle_lib
from oracle_lib import golden

pytestmark = pytest.mark.gpu
O = oracle_lib.oracle()


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d, _lib
    return torch, hfp3d, _lib


def load_mesh(stem):
    nods = np.load(golden(f"Nodes_{stem}.npy"))
    conn = np.load(golden(f"Elems_{stem}.npy")).astype(np.int64)
    conn[:3] -= 1
    return nods, conn


def stresses(torch, hx, lib, entry, dm, G, nu, dd, pts, s_inf):
    ctx = hx.get_context()
    n_el, n_pts = dd.shape[0] // 6, pts.shape[1]
    with torch.cuda.stream(ctx.stream):
        d_dd = torch.from_numpy(np.ascontiguousarray(dd.T)).to(ctx.device)
        d_pts = torch.from_numpy(np.ascontiguousarray(pts.T)).to(ctx.device)
        out = torch.full((6, n_pts), np.nan, dtype=torch.float64, device=ctx.device)
    ctx.stream.synchronize()
    s = np.ascontiguousarray(s_inf, dtype=np.float64)
    rc = getattr(ctx.lib, entry)(ctx.handle, dm.handle, G, nu, hx._ptr(d_dd), 6 * n_el, hx._ptr(d_pts), n_pts,
                                 hx._np_ptr(s), hx._ptr(out), n_pts)
    ctx.check(rc)
    ctx.stream.synchronize()
    return out.t().cpu().numpy()


@pytest.mark.parametrize("order", ["noalloc first", "allocating first"])
def test_noalloc_stress_matches_allocating_entry(env, order):
    torch, hx, lib = env
    rng = np.random.default_rng(4)
    cases = [("pennymesh121el_32", 1 / 2.6, 0.3, 3000), ("pennymesh24el_32", 1.0, 0.35, 700)]  # second smaller
    for stem, G, nu, n_pts in cases:
        print(f"case: stress_noalloc {stem} {n_pts} points ({order})", flush=True)
        nods, conn = load_mesh(stem)
        n_el = conn.shape[1]
        dm = hx.DeviceMesh(hx.Mesh_Geom_T(nods=nods, conn=conn), None, 0.0)
        dd = rng.normal(size=(6 * n_el, 3))
        pts = rng.uniform(-1.5, 1.5, size=(3, n_pts))
        s_inf = rng.normal(size=6)
        entries = ["hfpx_stress_noalloc", "hfpx_stress"]
        if order != "noalloc first":
            entries.reverse()
        got = {e: stresses(torch, hx, lib, e, dm, G, nu, dd, pts, s_inf) for e in entries}
        assert np.isfinite(got["hfpx_stress_noalloc"]).all()
        assert np.array_equal(got["hfpx_stress_noalloc"], got["hfpx_stress"])
        So = O.stress(nods, conn, G, nu, dd, pts, s_inf, nthreads=16)
        assert np.abs(got["hfpx_stress_noalloc"] - So).max() / np.abs(So).max() < 1e-12


def test_noalloc_stress_rejects_beyond_workspace(env):
    """18 n_el + 6 n_pts above the 8,388,608-double workspace: refused with
    HFPX_ERR_INVALID before any launch (the buffers handed in are real ones
    of the stated size)."""
    torch, hx, lib = env
    print("case: stress_noalloc beyond its workspace", flush=True)
    nods, conn = load_mesh("pennymesh24el_32")
    n_el = conn.shape[1]
    n_pts = (8 * 2 ** 20 - 18 * n_el) // 6 + 1
    ctx = hx.get_context()
    dm = hx.DeviceMesh(hx.Mesh_Geom_T(nods=nods, conn=conn), None, 0.0)
    with torch.cuda.stream(ctx.stream):
        d_dd = torch.zeros((3, 6 * n_el), dtype=torch.float64, device=ctx.device)
        d_pts = torch.zeros((n_pts, 3), dtype=torch.float64, device=ctx.device)
        out = torch.zeros((6, n_pts), dtype=torch.float64, device=ctx.device)
    ctx.stream.synchronize()
    s = np.zeros(6)
    rc = ctx.lib.hfpx_stress_noalloc(ctx.handle, dm.handle, 1.0, 0.3, hx._ptr(d_dd), 6 * n_el, hx._ptr(d_pts), n_pts,
                                     hx._np_ptr(s), hx._ptr(out), n_pts)
    assert rc == lib.HFPX_ERR_INVALID
    assert b"scratch" in ctx.lib.hfpx_last_error(ctx.handle)
    # one point fewer fits
    rc = ctx.lib.hfpx_stress_noalloc(ctx.handle, dm.handle, 1.0, 0.3, hx._ptr(d_dd), 6 * n_el, hx._ptr(d_pts),
                                     n_pts - 1, hx._np_ptr(s), hx._ptr(out), n_pts)
    ctx.check(rc)
    ctx.stream.synchronize()
