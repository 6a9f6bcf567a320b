"""Whole-pipeline parity on the benchmark configurations that fit a test:
the 1,025-element penny (FP64 GPU drop-in) at full size, and the non-planar
wavy crack (FP32 LU + FP64 GMRES-IR, stresses off the crack plane) at a
reduced element count.  Both against the oracle's FP64 matrix, LAPACK
solutions of it and the oracle's stress evaluation; the penny also against
Sneddon's opening."""
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
    from hfpx3d_amd import hfp3d
    return torch, hfp3d


def node_xyz(nods, conn):
    """Coordinates of the rows of the nodal DD array (row 6 e + node:
    vertices, then the mid-edge node opposite each vertex)."""
    xyz = np.zeros((6 * conn.shape[1], 3))
    for e in range(conn.shape[1]):
        v = nods[:3, conn[:3, e]].T
        for q in range(3):
            xyz[6 * e + q] = v[q]
            xyz[6 * e + 3 + q] = 0.5 * (v[(q + 1) % 3] + v[(q + 2) % 3])
    return xyz


def test_c2_penny1025_fp64_pipeline(env):
    """1,025 elements (18,450 DoF), E = 1, nu = 0.25, p = 1: the FP64 GPU
    matrix against the oracle's, the GPU FP64 LU solution against LAPACK on
    the oracle matrix (both to 1e-12 relative), and the opening against
    Sneddon's profile (0.5 % of its peak for r < 0.9; the tip elements, which
    cannot carry the square-root profile, to 8 %)."""
    torch, hx = env
    sl = pytest.importorskip("scipy.linalg")
    print("case: C2 penny1025 FP64", flush=True)
    nods = np.load(golden("Nodes_pennymesh1025el_32.npy"))
    conn = np.load(golden("Elems_pennymesh1025el_32.npy")).astype(np.int64)
    conn[:3] -= 1
    G, nu, p = 1 / 2.5, 0.25, 1.0
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    md = hx.Mesh_Data_T(mesh=mesh)
    npar = hx.Num_Param_T(beta=0.125, tip_type=0)
    load = hx.Load_T(s_inf=np.array([0, 0, p, 0, 0, 0.0]))
    sae = hx.SAE_T()
    sae.matrix = hx.make_3dbem_matrix_s(G, nu, mesh, npar, md.dof_h_dd)
    Ag = sae.matrix.cpu().numpy()
    Ao, dof = O.assemble(nods, conn, G, nu, 0.125, 0, nthreads=16)
    assert np.abs(Ag - Ao).max() / np.abs(Ao).max() < 1e-12
    del Ag
    hx.add_s_inf_to_3dbem_rhs(md, load, sae)
    x = hx.LU(sae.matrix).solve(sae.rhs_v).cpu().numpy()
    b = O.rhs(nods, conn, dof, load.s_inf)
    xr = sl.solve(Ao, b, check_finite=False, overwrite_a=True)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-12
    dd = hx.write_dd_vector_to_md(x, md.dof_h_dd)
    xyz = node_xyz(nods, conn)
    r = np.hypot(xyz[:, 0], xyz[:, 1])
    exact = 4 * (1 - nu) * p / (np.pi * G) * np.sqrt(np.clip(1 - r ** 2, 0, None))
    err = np.abs(dd[:, 2] - exact)
    assert err[r < 0.9].max() < 5e-3 * exact.max()
    assert err.max() < 8e-2 * exact.max()
    assert np.abs(dd[:, :2]).max() < 1e-12 * exact.max()


def test_c5_wavy_reduced_mixed_pipeline(env):
    """The wavy crack z = 0.1 sin(2 pi x) sin(2 pi y) over the unit disk with
    400 elements (7,200 DoF), E = 1, nu = 0.3, p = 1 plus s_zz = -0.5
    (s_inf = diag(1, 1, 0.5)): FP64 matrix against the oracle's; FP32 LU +
    FP64 GMRES-IR solution against LAPACK on the oracle matrix (1e-10); and
    stresses at 500 points uniform in [-2, 2]^3 (seed 3) against the oracle's
    evaluation with the same DD (norm-wise 1e-11)."""
    torch, hx = env
    sl = pytest.importorskip("scipy.linalg")
    from hfpx3d_amd import meshgen
    print("case: C5 wavy400 mixed", flush=True)
    nods, conn = meshgen.wavy_disk(400)
    G, nu = 1 / 2.6, 0.3
    s_inf = np.array([1.0, 1.0, 0.5, 0, 0, 0])
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    md = hx.Mesh_Data_T(mesh=mesh)
    npar = hx.Num_Param_T(beta=0.125, tip_type=0)
    load = hx.Load_T(s_inf=s_inf)
    sae = hx.SAE_T()
    sae.matrix = hx.make_3dbem_matrix_s(G, nu, mesh, npar, md.dof_h_dd)
    Ao, dof = O.assemble(nods, conn, G, nu, 0.125, 0, nthreads=16)
    assert np.abs(sae.matrix.cpu().numpy() - Ao).max() / np.abs(Ao).max() < 1e-12
    hx.add_s_inf_to_3dbem_rhs(md, load, sae)
    X, rep = hx.solve_mixed(sae.matrix, sae.rhs_v, tol=1e-15)
    assert rep["converged"], rep
    x = X.cpu().numpy()
    xr = sl.solve(Ao, O.rhs(nods, conn, dof, s_inf), check_finite=False)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-10
    md.dd = hx.write_dd_vector_to_md(x, md.dof_h_dd)
    pts = np.random.default_rng(3).uniform(-2, 2, size=(3, 500))
    S = hx.make_3dbem_stress_f_s(G, nu, md, load, npar, pts)
    So = O.stress(nods, conn, G, nu, md.dd, pts, s_inf, nthreads=16)
    assert np.abs(S - So).max() / np.abs(So).max() < 1e-11
