"""HIP assembly / right-hand side through the C ABI against the oracle and
the reference fixtures (GPU)."""
import numpy as np
import pytest

import oracle_lib
from oracle_lib import golden

pytestmark = pytest.mark.gpu

O = oracle_lib.oracle()


def load_mesh(stem):
    nods = np.load(golden(f"Nodes_{stem}.npy"))
    conn = np.load(golden(f"Elems_{stem}.npy")).astype(np.int64)
    conn[:3] -= 1
    return nods, conn


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def hx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d
    return hfp3d


def gpu_matrix(hx, nods, conn, G, nu, tip=0, beta=0.125, dtype=None):
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    npar = hx.Num_Param_T(beta=beta, tip_type=tip)
    dof = hx.DoF_Handle_T()
    A = hx.make_3dbem_matrix_s(G, nu, mesh, npar, dof, dtype=dtype)
    return A.cpu().numpy(), dof


TRI = np.array([[0, 0.1, 0], [1.8, 0, 0], [1.2, 1.8, 0]]).T
BN = np.array([[0, 0.1, 0], [1.8, 0, 0], [1.2, 1.8, 0], [2.2, 1.72, -1.21]]).T
BC = np.array([[0, 1, 2], [2, 1, 3]]).T


def test_one_element_matches_matlab_and_oracle(hx):
    print("case: one element", flush=True)
    A, _ = gpu_matrix(hx, TRI, np.array([[0, 1, 2]]).T, 1.0, 0.35)
    Ao, _ = O.assemble(TRI, np.array([[0, 1, 2]]).T, 1.0, 0.35)
    assert rel(A, Ao) < 1e-13
    assert rel(A, np.load(golden("matlab_assembly_triangle.npy"))) < 2e-8


def test_bitriangle_matches_oracle(hx):
    print("case: bitriangle", flush=True)
    A, _ = gpu_matrix(hx, BN, BC, 1.0, 0.35)
    Ao, _ = O.assemble(BN, BC, 1.0, 0.35)
    assert rel(A, Ao) < 1e-13


@pytest.mark.parametrize("tip", [0, 1, 2])
def test_penny24_matches_oracle_and_matlab(hx, tip):
    print(f"case: penny24 tip {tip}", flush=True)
    nods, conn = load_mesh("pennymesh24el_32")
    A, dof = gpu_matrix(hx, nods, conn, 1.0, 0.35, tip=tip)
    Ao, dof_o = O.assemble(nods, conn, 1.0, 0.35, tip_type=tip)
    assert np.array_equal(dof.dof_h, dof_o)
    assert rel(A, Ao) < 1e-13
    if tip == 0:
        assert rel(A, np.load(golden("matlab_assembly_24el.npy"))) < 1e-8


@pytest.mark.parametrize("case", ["ellipse24rot", "wavy24"])
def test_synthetic_meshes_match_oracle(hx, case):
    print(f"case: {case}", flush=True)
    d = np.load(golden("ref_assembly.npz"))
    nods, conn = d[case + "_nods"], d[case + "_conn"]
    A, _ = gpu_matrix(hx, nods, conn, 1 / 2.6, 0.3)
    Ao, _ = O.assemble(nods, conn, 1 / 2.6, 0.3)
    assert rel(A, Ao) < 1e-13


def test_penny121_full_matches_oracle(hx):
    print("case: penny121", flush=True)
    nods, conn = load_mesh("pennymesh121el_32")
    A, _ = gpu_matrix(hx, nods, conn, 1 / 2.6, 0.3)
    Ao, _ = O.assemble(nods, conn, 1 / 2.6, 0.3)
    assert rel(A, Ao) < 1e-13


def test_penny1025_sampled_rows_match_oracle_and_reference(hx):
    """C2 mesh (18,450 DoF, G = 0.4, nu = 0.25), the committed reference rows
    of three target elements: the GPU rows within 1e-13 of the oracle's, and
    within 1e-12 of the reference's at every (collocation point, source
    element) pair except those where the oracle differs from the reference
    too.  Each of those lies on the extension of a source edge line, where
    the reference's formulas lose accuracy (the full-matrix adjudication by
    quadrature and 30-digit evaluation is in test_oracle.py)."""
    import ref_adjudicate as ra
    print("case: penny1025 rows", flush=True)
    nods, conn = load_mesh("pennymesh1025el_32")
    d = np.load(golden("ref_rows_pennymesh1025el_32.npz"))
    A, _ = gpu_matrix(hx, nods, conn, 0.4, 0.25)
    mx = np.abs(A).max()
    n_explained = 0
    for i, t in enumerate(d["targets"]):
        t = int(t)
        Ao = O.assemble_rows(nods, conn, 0.4, 0.25, t, t + 1)
        rows = slice(18 * t, 18 * t + 18)
        assert np.abs(A[rows] - Ao).max() / mx < 1e-13
        Rr = d["rows"][18 * i:18 * i + 18]
        pair = lambda M: np.abs(M - Rr).reshape(6, 3, -1, 18).max(axis=(1, 3)) / mx  # noqa: E731
        bad_gpu, bad_oracle = pair(A[rows]) > 1e-12, pair(Ao) > 1e-12
        assert not (bad_gpu & ~bad_oracle).any()
        for k, se in np.argwhere(bad_gpu):
            _, _, h, z, tau, _ = ra.pair_frame(nods, conn, t, k, se)
            ok, where = ra.on_edge_extension(tau, h, z)
            assert ok, (t, k, se, where)
            n_explained += 1
    print(f"pairs off the reference by > 1e-12, all on edge-line extensions: {n_explained}")


def test_fp32_copy_is_rounded_fp64(hx):
    import torch
    print("case: fp32 copy", flush=True)
    nods, conn = load_mesh("pennymesh121el_32")
    A64, _ = gpu_matrix(hx, nods, conn, 1.0, 0.3)
    A32, _ = gpu_matrix(hx, nods, conn, 1.0, 0.3, dtype=torch.float32)
    assert np.array_equal(A32, A64.astype(np.float32))


def test_rotated_planar_mesh_is_rotation_of_planar(hx):
    """The h-noise of a rotated planar crack does not change the operator:
    A(Q mesh) = (I x Q) A(mesh) (I x Q^T) to rounding."""
    print("case: rotation", flush=True)
    from hfpx3d_amd import meshgen
    nods, conn = meshgen.ellipse(300, 1.0, 0.5)
    Q = meshgen.rotation([1, 0, 0], np.pi / 6)
    rn = nods.copy()
    rn[:3] = Q @ nods[:3]
    A0, _ = gpu_matrix(hx, nods, conn, 1 / 2.6, 0.3)
    A1, _ = gpu_matrix(hx, rn, conn, 1 / 2.6, 0.3)
    T = np.kron(np.eye(A0.shape[0] // 3), Q)
    assert rel(A1, T @ A0 @ T.T) < 1e-12


@pytest.mark.parametrize("tip", [0, 1])
def test_rhs_matches_oracle_and_reference(hx, tip):
    print(f"case: rhs tip {tip}", flush=True)
    nods, conn = load_mesh("pennymesh24el_32")
    d = np.load(golden("ref_rhs_stress.npz"))
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    md = hx.Mesh_Data_T(mesh=mesh, dof_h_dd=hx.make_dof_h_crack(mesh, 2, tip))
    sae = hx.SAE_T()
    r = hx.add_s_inf_to_3dbem_rhs(md, hx.Load_T(s_inf=d["s_inf"]), sae).cpu().numpy()
    assert np.abs(r - d[f"rhs_t{tip}"]).max() < 1e-14
    ro = O.rhs(nods, conn, md.dof_h_dd.dof_h, d["s_inf"])
    assert np.abs(r - ro).max() < 1e-14


def test_rhs_many_load_cases(hx):
    print("case: rhs 64", flush=True)
    nods, conn = load_mesh("pennymesh121el_32")
    rng = np.random.default_rng(1)
    S = rng.normal(0, 0.2, size=(6, 64))
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    md = hx.Mesh_Data_T(mesh=mesh, dof_h_dd=hx.make_dof_h_crack(mesh, 2, 0))
    R = hx.add_s_inf_to_3dbem_rhs(md, hx.Load_T(s_inf=S), hx.SAE_T()).cpu().numpy()
    for k in (0, 17, 63):
        ro = O.rhs(nods, conn, md.dof_h_dd.dof_h, S[:, k])
        assert np.abs(R[:, k] - ro).max() < 1e-14
