"""Volume-control systems through the C ABI (hfpx_assemble_vc,
hfpx_vc_truncate) against the oracle (GPU).  The oracle's volume-control
functions are pinned in test_oracle_vc.py."""
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


def oracle_vc(nods, conn, G, nu, tip):
    A, dof = O.assemble(nods, conn, G, nu, tip_type=tip)
    row, col = O.vc_border(nods, conn, dof)
    n = A.shape[0]
    Avc = np.zeros((n + 1, n + 1), order="F")
    Avc[:n, :n], Avc[n, :n], Avc[:n, n] = A, row, col
    return Avc, dof


def gpu_vc(hx, nods, conn, G, nu, tip):
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    dof = hx.DoF_Handle_T()
    A64, A32 = hx.make_3dbem_matrix_vc(G, nu, mesh, hx.Num_Param_T(tip_type=tip), dof, with_f32=True)
    return A64, A32, dof


@pytest.mark.parametrize("stem,tip", [("pennymesh24el_32", 0), ("pennymesh24el_32", 1), ("pennymesh121el_32", 1),
                                      ("elliptical_440_el_32", 1)])
def test_vc_matrix_matches_oracle(hx, stem, tip):
    print(f"case: vc matrix {stem} tip {tip}", flush=True)
    nods, conn = load_mesh(stem)
    A64, A32, dof = gpu_vc(hx, nods, conn, 1.0, 0.35, tip)
    A, A32 = A64.cpu().numpy(), A32.cpu().numpy()
    Ao, dof_o = oracle_vc(nods, conn, 1.0, 0.35, tip)
    assert np.array_equal(dof.dof_h, dof_o)
    n = Ao.shape[0] - 1
    assert rel(A[:n, :n], Ao[:n, :n]) < 1e-13
    assert np.abs(A[n, :n] - Ao[n, :n]).max() <= 1e-13 * np.abs(Ao[n, :n]).max()
    assert np.array_equal(A[:n, n], Ao[:n, n])
    assert A[n, n] == 0.0
    assert np.array_equal(A32, A.astype(np.float32))


def test_vc_matrix_with_padded_lda(hx):
    """The C ABI with lda > n_dof + 1 leaves the padding rows alone."""
    import torch
    print("case: vc matrix padded lda", flush=True)
    nods, conn = load_mesh("pennymesh24el_32")
    ctx = hx.get_context()
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    dm = hx.DeviceMesh(mesh, hx.make_dof_h_crack(mesh, 2, 1), 0.125, ctx)
    n1 = dm.n_dof + 1
    A, lda = hx.alloc_colmajor(n1, n1, torch.float64, ctx, lda=n1 + 37)
    with torch.cuda.stream(ctx.stream):
        A.as_strided((lda, n1), (1, lda)).fill_(7.0)
    ctx.check(ctx.lib.hfpx_assemble_vc(ctx.handle, dm.handle, 1.0, 0.35, hx._ptr(A), lda, None, 0))
    ctx.stream.synchronize()
    full = A.as_strided((lda, n1), (1, lda)).cpu().numpy()
    Ao, _ = oracle_vc(nods, conn, 1.0, 0.35, 1)
    assert rel(full[:n1], Ao) < 1e-13
    assert (full[n1:] == 7.0).all()


@pytest.mark.parametrize("tsize", ["full", "orig", "used"])
def test_vc_truncate_matches_oracle(hx, tsize):
    """Original handle without fixed DoF (tip_type 0), truncated to the
    tip_type 1 handle: bitwise equal to the oracle on the same matrix."""
    print(f"case: vc truncate penny121 delta_t {tsize}", flush=True)
    nods, conn = load_mesh("pennymesh121el_32")
    A64, _, dof0 = gpu_vc(hx, nods, conn, 1.0, 0.35, 0)
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    dof1 = hx.make_dof_h_crack(mesh, 2, 1)
    rng = np.random.default_rng(7)
    m = {"full": dof0.dof_h.size, "orig": dof0.n_dof, "used": dof1.n_dof}[tsize]
    if tsize == "orig":
        # orig n_dof equals n_el * 18 here; the slot branch comes first in the
        # reference and takes the same values (DoF = slot for tip_type 0)
        assert m == dof0.dof_h.size
    dt = rng.normal(size=m)
    sae = hx.mod_3dbem_system_vc(A64, dof0, dof1, dt, 0.75, with_f32=True)
    M, M32, rhs = sae.matrix.cpu().numpy(), sae.matrix32.cpu().numpy(), sae.rhs_v.cpu().numpy()
    Mo, ro = O.mod_system_vc(A64.cpu().numpy(), dof0.dof_h, dof1.dof_h, dof1.n_dof, dt, 0.75)
    assert sae.n_dof == dof1.n_dof
    assert np.array_equal(M, Mo)
    assert np.array_equal(M32, Mo.astype(np.float32))
    assert np.array_equal(rhs, ro)


def test_vc_truncate_from_fixed_original_and_merged_dof(hx):
    """Original handle with fixed tip DoF (border at orig n_dof) and a used
    handle that maps two slots to one DoF: sums in the reference's order."""
    print("case: vc truncate merged", flush=True)
    nods, conn = load_mesh("pennymesh121el_32")
    A64, _, dof1 = gpu_vc(hx, nods, conn, 1.0, 0.35, 1)
    dh = dof1.dof_h.copy()
    last = dof1.n_dof - 1
    dh[dh == last] = 11
    dh[dh == last - 1] = 11
    used = hx.DoF_Handle_T(n_dof=dof1.n_dof - 2, dof_h=dh)
    dt = np.arange(used.n_dof, dtype=float)
    sae = hx.mod_3dbem_system_vc(A64, dof1, used, dt, -2.0)
    Mo, ro = O.mod_system_vc(A64.cpu().numpy(), dof1.dof_h, dh, used.n_dof, dt, -2.0)
    assert np.array_equal(sae.matrix.cpu().numpy(), Mo)
    assert np.array_equal(sae.rhs_v.cpu().numpy(), ro)


def test_vc_truncate_unreferenced_dof(hx):
    """A used DoF that no slot maps to gets a zero row, column and rhs entry
    (the general gather path; the oracle zero-fills the same way)."""
    print("case: vc truncate unreferenced DoF", flush=True)
    nods, conn = load_mesh("pennymesh24el_32")
    A64, _, dof0 = gpu_vc(hx, nods, conn, 1.0, 0.35, 0)
    dof1 = hx.make_dof_h_crack(hx.Mesh_Geom_T(nods=nods, conn=conn), 2, 1)
    used = hx.DoF_Handle_T(n_dof=dof1.n_dof + 1, dof_h=dof1.dof_h)
    dt = np.ones(dof1.dof_h.size)
    sae = hx.mod_3dbem_system_vc(A64, dof0, used, dt, 3.0, with_f32=True)
    Mo, ro = O.mod_system_vc(A64.cpu().numpy(), dof0.dof_h, dof1.dof_h, used.n_dof, dt, 3.0)
    M = sae.matrix.cpu().numpy()
    assert np.array_equal(M, Mo)
    assert np.array_equal(sae.matrix32.cpu().numpy(), Mo.astype(np.float32))
    assert np.array_equal(sae.rhs_v.cpu().numpy(), ro)
    assert (M[dof1.n_dof] == 0).all() and (M[:, dof1.n_dof] == 0).all()


def test_vc_truncate_rejects_invalid(hx):
    print("case: vc truncate invalid", flush=True)
    nods, conn = load_mesh("pennymesh24el_32")
    A64, _, dof1 = gpu_vc(hx, nods, conn, 1.0, 0.35, 1)
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    dof0 = hx.make_dof_h_crack(mesh, 2, 0)
    with pytest.raises(Exception, match="fixed in the original"):
        hx.mod_3dbem_system_vc(A64, dof1, hx.DoF_Handle_T(n_dof=dof1.n_dof, dof_h=dof0.dof_h.clip(max=dof1.n_dof - 1)),
                               np.zeros(dof1.n_dof), 0.0)
    with pytest.raises(Exception, match="delta_t has"):
        hx.mod_3dbem_system_vc(A64, dof1, dof1, np.zeros(dof1.n_dof + 5), 0.0)


def test_vc_solve_prescribed_volume(hx):
    """Penny (1025 elements, tip DoF fixed): the volume-control system with
    zero traction change and volume increment dV, solved with the FP64 LU and
    with the FP32 LU + FP64 refinement.  The two agree to 1e-10, the volume
    constraint holds, and d = p d1 with p = dV / (v.d1), d1 the solution for
    the far-field load sigma_zz = 1 (right-hand side = e3 = -column n)."""
    import torch
    print("case: vc solve penny1025", flush=True)
    nods, conn = load_mesh("pennymesh1025el_32")
    A64, A32, dof = gpu_vc(hx, nods, conn, 1.0, 0.35, 1)
    n = dof.n_dof
    dv = 0.125
    sae = hx.mod_3dbem_system_vc(A64, dof, dof, np.zeros(n), dv, with_f32=True)
    assert torch.equal(sae.matrix, A64)
    x64 = hx.LU(sae.matrix.clone()).solve(sae.rhs_v).cpu().numpy()
    xm, rep = hx.solve_mixed(sae.matrix, sae.rhs_v, LU32=hx.LU(sae.matrix32))
    xm = xm.cpu().numpy()
    assert rep["converged"]
    assert np.abs(xm - x64).max() <= 1e-10 * np.abs(x64).max()
    A = A64.cpu().numpy()
    assert abs(A[n, :n] @ x64[:n] - dv) < 1e-10 * dv
    # the pressure solve on the same n x n block
    Ablk = torch.from_numpy(np.asfortranarray(A[:n, :n])).to(A64.device)
    Ablk = Ablk.t().contiguous().t()
    d1 = hx.LU(Ablk).solve(torch.from_numpy(-A[:n, n].copy()).to(A64.device)).cpu().numpy()
    p = dv / (A[n, :n] @ d1)
    assert abs(x64[n] - p) < 1e-9 * abs(p)
    assert np.abs(x64[:n] - p * d1).max() < 1e-9 * np.abs(x64[:n]).max()
    # Sneddon's volume per unit pressure, 8 (1 - nu) a^3 / (3 G), within the
    # discretisation error of the 1025-element mesh
    assert abs((A[n, :n] @ d1) / (8 * 0.65 / 3) - 1) < 0.05
