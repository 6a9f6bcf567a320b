"""hfpx_mesh_create_npy: the reference's npy mesh files straight to a
device mesh (loader, index origin, DoF handle, upload) against the Python
path on the same files, and its error reporting (GPU)."""
import ctypes
import os

import numpy as np
import pytest

from oracle_lib import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d
    return hfp3d


@pytest.mark.parametrize("stem,tip", [("pennymesh121el_32", 1), ("pennymesh24el_32", 2), ("pennymesh24el_32", 0)])
def test_mesh_from_npy_matches_python_path(hx, stem, tip):
    import torch
    print(f"case: mesh from npy {stem} tip {tip}", flush=True)
    from hfpx3d_amd import _lib
    ctx = hx.get_context()
    lib = ctx.lib
    h = ctypes.c_void_p()
    nodes, elems = golden(f"Nodes_{stem}.npy"), golden(f"Elems_{stem}.npy")
    ctx.check(lib.hfpx_mesh_create_npy(ctx.handle, nodes.encode(), elems.encode(), 1, tip, 0.125, ctypes.byref(h)))
    n = int(lib.hfpx_mesh_n_dof(h))
    A, lda = hx.alloc_colmajor(n, n, torch.float64, ctx)
    ctx.enter()
    ctx.check(lib.hfpx_assemble(ctx.handle, h, 1.0, 0.35, ctypes.c_void_p(A.data_ptr()), lda))
    ctx.leave()
    lib.hfpx_mesh_destroy(h)
    # the Python path on the same files
    nods = np.load(nodes)
    conn = np.load(elems).astype(np.int64)
    conn[:3] -= 1
    dof = hx.DoF_Handle_T()
    Ap = hx.make_3dbem_matrix_s(1.0, 0.35, hx.Mesh_Geom_T(nods=nods, conn=conn), hx.Num_Param_T(tip_type=tip), dof)
    assert dof.n_dof == n
    assert torch.equal(A.cpu(), Ap.cpu())
    assert _lib.HFPX_OK == 0


def test_mesh_from_npy_reports_missing_file(hx, tmp_path):
    from hfpx3d_amd import _lib
    print("case: mesh from npy, missing file", flush=True)
    ctx = hx.get_context()
    h = ctypes.c_void_p()
    missing = os.path.join(str(tmp_path), "nope.npy")
    rc = ctx.lib.hfpx_mesh_create_npy(ctx.handle, missing.encode(), golden("Elems_pennymesh24el_32.npy").encode(), 1,
                                      0, 0.125, ctypes.byref(h))
    assert rc == _lib.HFPX_ERR_INVALID
    assert b"nope.npy" in ctx.lib.hfpx_last_error(ctx.handle)
    rc = ctx.lib.hfpx_mesh_create_npy(ctx.handle, golden("Nodes_pennymesh24el_32.npy").encode(),
                                      golden("Elems_pennymesh24el_32.npy").encode(), 1, 3, 0.125, ctypes.byref(h))
    assert rc == _lib.HFPX_ERR_INVALID
