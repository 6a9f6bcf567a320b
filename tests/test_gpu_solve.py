"""Mixed-precision solve (FP32 LU + FP64 iterative refinement, classic and
GMRES steps): on BEM matrices with many load cases against LAPACK on the
oracle's matrix (and against the FP64
GPU LU, itself pinned to LAPACK in test_gpu_configs.py); and its failure
handling: refinement that cannot converge falls back to an FP64 LU of a
copy of the matrix (fell_back = 1), a non-finite right-hand side is
reported as not converged with a NaN backward error, and a singular FP32
factor is recovered from."""
import numpy as np
import pytest

import oracle_lib
from oracle_lib import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d
    return torch, hfp3d


def sinf_cases(nrhs, seed=1):
    """column 0: pressure p = 1 (s_inf = I); others: symmetric N(0, 0.2^2)."""
    rng = np.random.default_rng(seed)
    S = np.zeros((6, nrhs))
    S[:3, 0] = 1.0
    S[:, 1:] = rng.normal(0.0, 0.2, size=(6, nrhs - 1))
    return S


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_gmres_ir_matches_lapack_on_oracle_matrix(env, mode):
    """penny121 (2,178 DoF), nu = 0.3, 64 load cases: the mixed solution
    against scipy's LAPACK solve of the oracle's matrix and right-hand
    sides, 1e-10 per load case, for every refinement mode (0: classic steps
    while they contract 100-fold, then GMRES; 1: GMRES only; 2: classic
    only)."""
    torch, hx = env
    from hfpx3d_amd import _lib
    sl = pytest.importorskip("scipy.linalg")
    print(f"case: gmres-ir penny121 vs LAPACK on the oracle matrix, IR mode {mode}", flush=True)
    O = oracle_lib.oracle()
    nods = np.load(golden("Nodes_pennymesh121el_32.npy"))
    conn = np.load(golden("Elems_pennymesh121el_32.npy")).astype(np.int64)
    conn[:3] -= 1
    G, nu = 1 / 2.6, 0.3
    S = sinf_cases(64)
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    md = hx.Mesh_Data_T(mesh=mesh)
    A = hx.make_3dbem_matrix_s(G, nu, mesh, hx.Num_Param_T(beta=0.125, tip_type=0), md.dof_h_dd)
    B = hx.add_s_inf_to_3dbem_rhs(md, hx.Load_T(s_inf=S), hx.SAE_T())
    ctx = hx.get_context()
    hx.tune_set(_lib.KNOB_IR_MODE, mode, ctx)
    try:
        X, rep = hx.solve_mixed(A, B, tol=1e-15)
    finally:
        hx.tune_set(_lib.KNOB_IR_MODE, 0, ctx)
    assert rep["converged"] and not rep["fell_back"], rep
    if mode == 2:
        assert rep["inner_iterations"] == 0, rep
    if mode == 1:
        assert rep["inner_iterations"] >= rep["iterations"], rep
    Ao, dof = O.assemble(nods, conn, G, nu, 0.125, 0)
    Bo = np.column_stack([O.rhs(nods, conn, dof, S[:, r]) for r in range(S.shape[1])])
    Xr = sl.solve(Ao, Bo)
    err = np.abs(X.cpu().numpy() - Xr).max(axis=0) / np.abs(Xr).max(axis=0)
    assert err.max() < 1e-10, err.max()


def _device(torch, hx, a):
    ctx = hx.get_context()
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    d, _ = hx.alloc_colmajor(a.shape[0], a.shape[1], torch.float64, ctx)
    with torch.cuda.stream(ctx.stream):
        d.copy_(torch.from_numpy(a))
    return d


def _ill_conditioned(n, cond, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.normal(size=(n, n)))
    V, _ = np.linalg.qr(rng.normal(size=(n, n)))
    s = np.logspace(0, -np.log10(cond), n)
    return (U * s) @ V.T


def _berr(A, X, B):
    r = np.abs(B - A @ X).max(axis=0)
    return (r / (np.abs(A).sum(axis=1).max() * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))).max()


def test_ir_switches_to_gmres_when_classic_steps_contract_weakly(env):
    """kappa(A) = 1e7: kappa eps_fp32 is near 1, so a classic step x += M^-1 r
    contracts the backward error far less than 100-fold; the default mode
    then continues with GMRES steps (inner_iterations > 0) and converges
    without the FP64 fallback, to LAPACK's solution within kappa eps_64."""
    torch, hx = env
    print("case: IR classic -> GMRES switch, kappa 1e7", flush=True)
    n = 600
    A = _ill_conditioned(n, 1e7, 13)
    B = np.random.default_rng(14).normal(size=(n, 4))
    X, rep = hx.solve_mixed(_device(torch, hx, A), _device(torch, hx, B), tol=1e-15, max_outer=10, gmres_m=30)
    assert rep["converged"] and not rep["fell_back"], rep
    assert rep["inner_iterations"] > 0 and rep["iterations"] >= 2, rep
    x = X.cpu().numpy()
    assert _berr(A, x, B) < 1e-15
    xr = np.linalg.solve(A, B)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-7


def test_gmres_ir_falls_back_to_fp64_lu_when_fp32_cannot_precondition(env):
    """kappa(A) = 1e13, far beyond 1 / eps_fp32: the FP32 factors do not
    precondition A, refinement stops (3 outer steps of GMRES(5) allowed),
    and the solution comes from the FP64 LU of a copy of A: fell_back = 1,
    backward error at FP64 level, A itself unchanged.  With the fallback
    switched off the same call reports HFPX_ERR_NOT_CONVERGED."""
    torch, hx = env
    from hfpx3d_amd import _lib
    print("case: gmres-ir fallback, kappa 1e13", flush=True)
    n = 600
    A = _ill_conditioned(n, 1e13, 11)
    B = np.random.default_rng(12).normal(size=(n, 3))
    Ad, Bd = _device(torch, hx, A), _device(torch, hx, B)
    X, rep = hx.solve_mixed(Ad, Bd, tol=1e-15, max_outer=3, gmres_m=5)
    assert rep["fell_back"], rep
    assert rep["converged"], rep
    x = X.cpu().numpy()
    assert _berr(A, x, B) < 1e-15
    assert np.array_equal(Ad.cpu().numpy(), A)
    xr = np.linalg.solve(A, B)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-2  # both backward stable; kappa eps_64 ~ 1e-3
    ctx = hx.get_context()
    hx.tune_set(_lib.KNOB_IR_FALLBACK, 2, ctx)
    try:
        X2, rep2 = hx.solve_mixed(Ad, Bd, tol=1e-15, max_outer=3, gmres_m=5)
    finally:
        hx.tune_set(_lib.KNOB_IR_FALLBACK, 0, ctx)
    assert not rep2["converged"] and not rep2["fell_back"] and rep2["status"] == _lib.HFPX_ERR_NOT_CONVERGED


def test_gmres_ir_recovers_from_singular_fp32_factor(env):
    """Column 5 scaled by 2^-170: its FP32 rounding is exactly zero, so the
    FP32 LU meets a zero pivot that FP64 does not (reported through Status,
    as il::LU does, LU.h:94-98).  The preconditioned iterate is then not
    finite and the FP64 fallback solves the system.  The power-of-two column
    scale is exact, so LAPACK's solution of the unscaled matrix, rescaled,
    is the reference."""
    torch, hx = env
    print("case: gmres-ir singular FP32 factor", flush=True)
    n = 300
    rng = np.random.default_rng(21)
    A0 = rng.normal(size=(n, n)) + 20 * np.eye(n)
    A = A0.copy()
    A[:, 5] *= 2.0 ** -170
    assert not A[:, 5].astype(np.float32).any() and A[:, 5].all()
    B = rng.normal(size=(n, 2))
    Ad, Bd = _device(torch, hx, A), _device(torch, hx, B)
    ctx = hx.get_context()
    A32, _ = hx.alloc_colmajor(n, n, torch.float32, ctx)
    A32.copy_(Ad)
    st = hx.Status()
    lu32 = hx.LU(A32, status=st)
    assert st.error == "MatrixSingular"
    X, rep = hx.solve_mixed(Ad, Bd, tol=1e-15, LU32=lu32)
    assert rep["converged"] and rep["fell_back"], rep
    x = X.cpu().numpy()
    assert _berr(A, x, B) < 1e-15
    y = x.copy()
    y[5] *= 2.0 ** -170  # back to the unknowns of the unscaled system
    yr = np.linalg.solve(A0, B)
    assert np.abs(y - yr).max() / np.abs(yr).max() < 1e-12


def test_gmres_ir_reports_non_finite_rhs(env):
    """A NaN in B: not converged, NaN backward error, no fallback (the data
    themselves are not finite)."""
    torch, hx = env
    from hfpx3d_amd import _lib
    print("case: gmres-ir NaN right-hand side", flush=True)
    n = 400
    rng = np.random.default_rng(31)
    A = rng.normal(size=(n, n)) + 30 * np.eye(n)
    B = rng.normal(size=(n, 2))
    B[17, 1] = np.nan
    X, rep = hx.solve_mixed(_device(torch, hx, A), _device(torch, hx, B), tol=1e-14)
    assert not rep["converged"] and not rep["fell_back"]
    assert rep["status"] == _lib.HFPX_ERR_NOT_CONVERGED
    assert not np.isfinite(rep["backward_error"])


@pytest.mark.parametrize("stem,nrhs", [("pennymesh121el_32", 64), ("pennymesh1025el_32", 8)])
def test_gmres_ir_matches_fp64_lu(env, stem, nrhs):
    torch, hx = env
    print(f"case: gmres-ir {stem} nrhs={nrhs}", flush=True)
    nods = np.load(golden(f"Nodes_{stem}.npy"))
    conn = np.load(golden(f"Elems_{stem}.npy")).astype(np.int64)
    conn[:3] -= 1
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    npar = hx.Num_Param_T(beta=0.125, tip_type=0)
    md = hx.Mesh_Data_T(mesh=mesh)
    A = hx.make_3dbem_matrix_s(1 / 2.6, 0.3, mesh, npar, md.dof_h_dd)
    sae = hx.SAE_T()
    B = hx.add_s_inf_to_3dbem_rhs(md, hx.Load_T(s_inf=sinf_cases(nrhs)), sae)
    X, rep = hx.solve_mixed(A, B, tol=1e-15, max_outer=10)
    assert rep["converged"], rep
    # FP64 reference: LU of a copy
    n = A.shape[0]
    A2, _ = hx.alloc_colmajor(n, n, torch.float64, hx.get_context())
    A2.copy_(A)
    X64 = hx.LU(A2).solve(B)
    x, x64 = X.cpu().numpy(), X64.cpu().numpy()
    err = np.abs(x - x64).max(axis=0) / np.abs(x64).max(axis=0)
    print("max rel err per rhs", err.max(), rep, flush=True)
    assert err.max() <= 1e-10


def test_gmres_ir_random_matrix(env):
    torch, hx = env
    print("case: gmres-ir random", flush=True)
    rng = np.random.default_rng(5)
    n = 1500
    A = rng.normal(size=(n, n)) + 40 * np.eye(n)
    ctx = hx.get_context()
    Ad, _ = hx.alloc_colmajor(n, n, torch.float64, ctx)
    Ad.copy_(torch.from_numpy(A))
    B = rng.normal(size=(n, 5))
    Bd, _ = hx.alloc_colmajor(n, 5, torch.float64, ctx)
    Bd.copy_(torch.from_numpy(B))
    X, rep = hx.solve_mixed(Ad, Bd, tol=1e-15)
    assert rep["converged"]
    x = X.cpu().numpy()
    xr = np.linalg.solve(A, B)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-12
