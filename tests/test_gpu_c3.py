"""The headline configuration: the elliptical crack with semi-axes 1 and 0.5,
rotated 30 degrees about x, E = 1, nu = 0.3, 64 load cases (p = 1 plus 63
random symmetric far-field stresses with entries N(0, 0.2^2), seed 1),
assembled in FP64 together with its FP32 rounding, factorised by the FP32
LU and solved by FP64 GMRES iterative refinement.

* Reduced size (500 elements, 9,000 DoF): every stage against the oracle
  and LAPACK on the oracle's matrix: the FP64 matrix (1e-12 of its largest
  entry), the FP32 copy (the exact rounding), the 64 right-hand sides and
  the 64 mixed-precision solutions (1e-10 relative per load case).
* Full size (4,000 elements, 72,000 DoF), through properties that do not
  need a CPU solve of that size: sampled row blocks of the FP64 matrix
  against the oracle, sampled columns of the FP32 copy against the rounding
  of the FP64 ones, the normwise backward error of every solution against
  the FP64 matrix (recomputed here with torch, not taken from the report),
  and the mixed solutions against the plain FP64 GPU LU of the same matrix
  (1e-10).
"""
import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu
O = oracle_lib.oracle()
E_MOD, NU = 1.0, 0.3
G = E_MOD / (2 * (1 + NU))


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d, _lib
    return torch, hfp3d, _lib


def c3_problem(n_el, nrhs=64, seed=1):
    """The rotated ellipse and its load cases (draw order: one N(0, 0.2^2)
    matrix of 6 x (nrhs - 1), components xx yy zz xy xz yz)."""
    from hfpx3d_amd import meshgen
    R = meshgen.rotation([1.0, 0.0, 0.0], np.pi / 6)
    nods, conn = meshgen.ellipse(n_el, 1.0, 0.5, R)
    S = np.zeros((6, nrhs))
    S[:3, 0] = 1.0  # pressure p = 1 on both faces <=> s_inf = p I
    S[:, 1:] = np.random.default_rng(seed).normal(0.0, 0.2, size=(6, nrhs - 1))
    return nods, conn, S


def assemble_mixed(torch, hx, nods, conn, S):
    ctx = hx.get_context()
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    dm = hx.DeviceMesh(mesh, None, 0.125, ctx)
    n = dm.n_dof
    A64, lda = hx.alloc_colmajor(n, n, torch.float64, ctx)
    A32, lda32 = hx.alloc_colmajor(n, n, torch.float32, ctx)
    ctx.check(ctx.lib.hfpx_assemble_mixed(ctx.handle, dm.handle, G, NU, hx._ptr(A64), lda, hx._ptr(A32), lda32))
    B, ldb = hx.alloc_colmajor(n, S.shape[1], torch.float64, ctx)
    with torch.cuda.stream(ctx.stream):
        B.zero_()
    Sf = np.asfortranarray(S)
    ctx.check(ctx.lib.hfpx_rhs_sinf(ctx.handle, dm.handle, hx._np_ptr(Sf), S.shape[1], hx._ptr(B), ldb))
    ctx.stream.synchronize()
    return ctx, dm, A64, A32, B


def mixed_solve(torch, hx, _lib, ctx, A64, A32, B, tol=1e-14):
    """FP32 LU of A32 in place + FP64 GMRES-IR against A64."""
    import ctypes
    n = A64.shape[0]
    with torch.cuda.stream(ctx.stream):
        ipiv = torch.empty(n, dtype=torch.int32, device=ctx.device)
    X, ldx = hx.alloc_colmajor(n, B.shape[1], torch.float64, ctx)
    info = ctypes.c_int64(0)
    ctx.check(ctx.lib.hfpx_getrf(ctx.handle, _lib.HFPX_F32, hx._ptr(A32), A32.stride(1), n, hx._ptr(ipiv),
                                 ctypes.byref(info)))
    rep = _lib.IRReport()
    rc = ctx.lib.hfpx_solve_ir(ctx.handle, hx._ptr(A64), A64.stride(1), hx._ptr(A32), A32.stride(1), hx._ptr(ipiv), n,
                               hx._ptr(B), B.stride(1), hx._ptr(X), ldx, B.shape[1], tol, 10, 20, ctypes.byref(rep))
    ctx.check(rc)
    ctx.stream.synchronize()
    return X, rep


def test_c3_reduced_against_oracle_and_lapack(env):
    torch, hx, _lib = env
    sl = pytest.importorskip("scipy.linalg")
    print("case: C3 rotated ellipse, 500 elements, 64 RHS", flush=True)
    nods, conn, S = c3_problem(500)
    ctx, dm, A64, A32, B = assemble_mixed(torch, hx, nods, conn, S)
    n = dm.n_dof
    assert n == 9000
    Ao, dof = O.assemble(nods, conn, G, NU, 0.125, 0, nthreads=16)
    Ag = A64.cpu().numpy()
    assert np.abs(Ag - Ao).max() / np.abs(Ao).max() < 1e-12
    assert np.array_equal(A32.cpu().numpy(), Ag.astype(np.float32))
    Bo = np.column_stack([O.rhs(nods, conn, dof, S[:, r]) for r in range(S.shape[1])])
    assert np.abs(B.cpu().numpy() - Bo).max() < 1e-14
    del Ag
    X, rep = mixed_solve(torch, hx, _lib, ctx, A64, A32, B)
    assert rep.converged and rep.backward_error <= 1e-14 and not rep.fell_back, \
        (rep.iterations, rep.backward_error, rep.fell_back)
    Xr = sl.solve(Ao, Bo, check_finite=False, overwrite_a=True)
    err = (np.abs(X.cpu().numpy() - Xr).max(axis=0) / np.abs(Xr).max(axis=0))
    print(f"C3 reduced: IR {rep.iterations} outer / {rep.inner_iterations} inner, max rel err {err.max():.2e}")
    assert err.max() < 1e-10, err.max()


def test_c3_full_size_properties(env):
    torch, hx, _lib = env
    import ctypes
    print("case: C3 rotated ellipse, 4,000 elements (72,000 DoF), 64 RHS", flush=True)
    nods, conn, S = c3_problem(4000)
    ctx, dm, A64, A32, B = assemble_mixed(torch, hx, nods, conn, S)
    n = dm.n_dof
    assert n == 72000
    # sampled row blocks (every column) of the FP64 matrix against the oracle
    rng = np.random.default_rng(5)
    targets = sorted(rng.choice(4000, size=6, replace=False).tolist())
    mx = 0.0
    errs = []
    for t in targets:
        rows = O.assemble_rows(nods, conn, G, NU, t, t + 1, nthreads=16)
        g = A64[18 * t:18 * t + 18].cpu().numpy()
        mx = max(mx, np.abs(rows).max())
        errs.append(np.abs(g - rows).max())
    assert max(errs) / mx < 1e-12, max(errs) / mx
    # sampled columns of the FP32 copy: the rounding of the FP64 ones
    cols = torch.from_numpy(rng.choice(n, size=256, replace=False)).to(ctx.device)
    with torch.cuda.stream(ctx.stream):
        same = bool((A32[:, cols] == A64[:, cols].to(torch.float32)).all().item())
    assert same
    # |A|_inf of the FP64 matrix, in column blocks (no full-size temporary)
    with torch.cuda.stream(ctx.stream):
        rs = torch.zeros(n, dtype=torch.float64, device=ctx.device)
        for c0 in range(0, n, 4096):
            rs += A64[:, c0:c0 + 4096].abs().sum(dim=1)
        anorm = rs.max().item()
        A2, _ = hx.alloc_colmajor(n, n, torch.float64, ctx)
        A2.copy_(A64)
        X64, _ = hx.alloc_colmajor(n, B.shape[1], torch.float64, ctx)
        X64.copy_(B)
    ctx.stream.synchronize()
    X, rep = mixed_solve(torch, hx, _lib, ctx, A64, A32, B)
    assert rep.converged and not rep.fell_back
    with torch.cuda.stream(ctx.stream):
        R = B - torch.matmul(A64, X.contiguous())
        berr = (R.abs().amax(dim=0) / (anorm * X.abs().amax(dim=0) + B.abs().amax(dim=0))).max().item()
    ctx.stream.synchronize()
    print(f"C3 72k: IR {rep.iterations} outer / {rep.inner_iterations} inner, backward error {berr:.2e} "
          f"(report {rep.backward_error:.2e})")
    assert berr <= 1e-14, berr
    # the plain FP64 GPU path on the same matrix
    with torch.cuda.stream(ctx.stream):
        ipiv = torch.empty(n, dtype=torch.int32, device=ctx.device)
    info = ctypes.c_int64(0)
    ctx.check(ctx.lib.hfpx_getrf(ctx.handle, _lib.HFPX_F64, hx._ptr(A2), n, n, hx._ptr(ipiv), ctypes.byref(info)))
    ctx.check(ctx.lib.hfpx_getrs(ctx.handle, _lib.HFPX_F64, hx._ptr(A2), n, hx._ptr(ipiv), n, hx._ptr(X64), n,
                                 B.shape[1]))
    with torch.cuda.stream(ctx.stream):
        err = ((X - X64).abs().amax(dim=0) / X64.abs().amax(dim=0)).max().item()
    ctx.stream.synchronize()
    print(f"C3 72k: mixed vs FP64 LU max rel err {err:.2e}")
    assert err < 1e-10, err
    del A2, X64, A32, A64
    torch.cuda.empty_cache()
