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


def _write_npy_bytes(path, header, payload=b""):
    """A .npy v1.0 file with the given header dict text and raw payload."""
    h = header.encode()
    pad = 64 - (10 + len(h) + 1) % 64
    with open(path, "wb") as f:
        f.write(b"\x93NUMPY\x01\x00" + (len(h) + pad + 1).to_bytes(2, "little") + h + b" " * pad + b"\n")
        f.write(payload)


@pytest.mark.parametrize("header,payload", [
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, 99999999999), }", b"\0" * 72),   # shape beyond the data
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, -5), }", b"\0" * 72),            # negative dimension
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, x), }", b"\0" * 72),             # not a number
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, 3", b"\0" * 72),                 # unterminated
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, 4), }", b"\0" * 40),             # truncated data
    ("{'descr': '<c16', 'fortran_order': False, 'shape': (3, 3), }", b"\0" * 144),           # unsupported dtype
])
def test_mesh_from_npy_rejects_corrupt_nodes(hx, tmp_path, header, payload):
    from hfpx3d_amd import _lib
    print(f"case: corrupt npy {header[-24:]}", flush=True)
    ctx = hx.get_context()
    bad = os.path.join(str(tmp_path), "bad.npy")
    _write_npy_bytes(bad, header, payload)
    h = ctypes.c_void_p()
    rc = ctx.lib.hfpx_mesh_create_npy(ctx.handle, bad.encode(), golden("Elems_pennymesh24el_32.npy").encode(), 1, 0,
                                      0.125, ctypes.byref(h))
    assert rc == _lib.HFPX_ERR_INVALID
    assert not h.value


def test_mesh_from_npy_rejects_out_of_range_vertex(hx, tmp_path):
    """A vertex id past the node table is refused before the DoF handle
    reads the node rows it names."""
    from hfpx3d_amd import _lib
    print("case: npy vertex id out of range", flush=True)
    ctx = hx.get_context()
    conn = np.load(golden("Elems_pennymesh24el_32.npy")).copy()
    conn[1, 4] = 10 ** 6
    path = os.path.join(str(tmp_path), "elems.npy")
    np.save(path, conn)
    h = ctypes.c_void_p()
    rc = ctx.lib.hfpx_mesh_create_npy(ctx.handle, golden("Nodes_pennymesh24el_32.npy").encode(), path.encode(), 1, 1,
                                      0.125, ctypes.byref(h))
    assert rc == _lib.HFPX_ERR_INVALID
    assert b"out of range" in ctx.lib.hfpx_last_error(ctx.handle)


@pytest.mark.parametrize("kind", ["collinear", "repeated vertex"])
def test_mesh_rejects_degenerate_element(hx, kind):
    """A zero-area element has no frame (its geometry would divide by
    zero): mesh creation refuses it instead of assembling inf / NaN."""
    from hfpx3d_amd import _lib
    print(f"case: degenerate element ({kind})", flush=True)
    nods = np.load(golden("Nodes_pennymesh24el_32.npy"))[:3].copy()
    conn = np.load(golden("Elems_pennymesh24el_32.npy")).astype(np.int64)[:3] - 1
    a, b, c = conn[:, 5]
    if kind == "collinear":
        nods[:, c] = 0.5 * (nods[:, a] + nods[:, b])
    else:
        conn[2, 5] = a
    with pytest.raises(_lib.HfpxError) as e:
        hx.DeviceMesh(hx.Mesh_Geom_T(nods=nods, conn=conn), None, 0.125)
    assert e.value.status == _lib.HFPX_ERR_INVALID and "degenerate" in str(e.value)
