"""Mixed-precision solve (FP32 LU + FP64 GMRES-IR) against the FP64 LU
solution (GPU), on a BEM matrix with many load cases."""
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
