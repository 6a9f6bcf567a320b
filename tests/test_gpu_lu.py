"""MFMA GEMM, LU with partial pivoting and triangular solves (GPU) against
LAPACK (scipy) and the oracle, through the C ABI."""
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


def to_dev(torch, hx, a, dtype):
    """numpy (n x m) -> column-major device tensor, copied on the context's
    stream (the tests that call the C ABI directly launch on that stream;
    a copy on torch's current stream would race with them)."""
    ctx = hx.get_context()
    t, _ = hx.alloc_colmajor(a.shape[0], a.shape[1], dtype, ctx)
    with torch.cuda.stream(ctx.stream):
        t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
    return t


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("M,N,K", [(128, 128, 16), (300, 257, 33), (1000, 64, 256), (77, 5, 3), (513, 700, 100),
                                   (3000, 64, 20000), (5, 3, 9000), (1000, 20, 10000), (2000, 1, 12000)])
def test_gemm_sub(env, dt, M, N, K):
    torch, hx = env
    print(f"case: gemm {dt} {M}x{N}x{K}", flush=True)
    dtype = torch.float64 if dt == "f64" else torch.float32
    rng = np.random.default_rng(M + N + K)
    A, B, C = rng.normal(size=(M, K)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    Cd = to_dev(torch, hx, C, dtype)
    hx.gemm_sub(Cd, to_dev(torch, hx, A, dtype), to_dev(torch, hx, B, dtype))
    got = Cd.cpu().numpy()
    ref = C - A @ B
    tol = 1e-13 if dt == "f64" else 2e-5
    assert np.abs(got - ref).max() / (np.abs(A).max() * np.abs(B).max() * K) < tol


@pytest.mark.parametrize("variant", range(1, 18))
@pytest.mark.parametrize("M,N,K", [(300, 257, 33), (513, 700, 100), (1000, 300, 256), (130, 129, 17),
                                   (256, 256, 32), (77, 200, 20), (512, 384, 256)])
def test_sgemm_variants(env, variant, M, N, K):
    """Every FP32 trailing-update kernel variant, on shapes that cut through
    its tiles and K chunks."""
    torch, hx = env
    from hfpx3d_amd import _lib
    print(f"case: sgemm variant {variant} {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(7 * M + N + K + variant)
    A, B, C = rng.normal(size=(M, K)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    Cd = to_dev(torch, hx, C, torch.float32)
    old = hx.tune_get(_lib.KNOB_GEMM_F32)
    hx.tune_set(_lib.KNOB_GEMM_F32, variant)
    try:
        hx.gemm_sub(Cd, to_dev(torch, hx, A, torch.float32), to_dev(torch, hx, B, torch.float32))
    finally:
        hx.tune_set(_lib.KNOB_GEMM_F32, old)
    got = Cd.cpu().numpy()
    ref = C.astype(np.float32).astype(np.float64) - A.astype(np.float32).astype(np.float64) @ \
        B.astype(np.float32).astype(np.float64)
    assert np.abs(got - ref).max() / (np.abs(A).max() * np.abs(B).max() * K) < 2e-6


@pytest.mark.parametrize("M,N,K", [(4500, 4200, 1024), (700, 650, 40), (4096, 4096, 512)])
def test_sgemm_pipelined_split_is_bitwise_variant_14(env, M, N, K):
    """Variant 15 issues the products of variant 14 in another order across
    accumulators but the same order within each: bitwise the same C."""
    torch, hx = env
    from hfpx3d_amd import _lib
    print(f"case: sgemm 15 vs 14 {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M + 3 * N + K)
    A, B, C = rng.normal(size=(M, K)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    Ad, Bd = to_dev(torch, hx, A, torch.float32), to_dev(torch, hx, B, torch.float32)
    out = {}
    old = hx.tune_get(_lib.KNOB_GEMM_F32)
    try:
        for v in (14, 15):
            hx.tune_set(_lib.KNOB_GEMM_F32, v)
            Cd = to_dev(torch, hx, C, torch.float32)
            hx.gemm_sub(Cd, Ad, Bd)
            out[v] = Cd.cpu().numpy()
    finally:
        hx.tune_set(_lib.KNOB_GEMM_F32, old)
    assert np.array_equal(out[14], out[15])


@pytest.mark.parametrize("M,N,K,ld_pad", [(4500, 4200, 1024, 0), (700, 650, 40, 0), (4096, 4096, 512, 0),
                                          (1000, 777, 203, 1), (300, 129, 64, 3)])
def test_sgemm_fused_split_is_bitwise_the_split_pass(env, M, N, K, ld_pad):
    """Variants 16 / 17 split the operands in the product's load stage; the
    parts and products are those of the split pass, so C is bitwise that of
    variant 14 (128 x 256 tiles) / 13 (128 x 128), also with leading
    dimensions that are not a multiple of 4 (B read with 4-byte loads)."""
    torch, hx = env
    from hfpx3d_amd import _lib
    print(f"case: sgemm fused split {M}x{N}x{K} pad {ld_pad}", flush=True)
    rng = np.random.default_rng(M + 5 * N + K + ld_pad)
    A, B, C = rng.normal(size=(M, K)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    lda, ldb, ldc = M + ld_pad, K + ld_pad, M + ld_pad
    Ad = to_dev(torch, hx, np.pad(A, ((0, ld_pad), (0, 0))), torch.float32)
    Bd = to_dev(torch, hx, np.pad(B, ((0, ld_pad), (0, 0))), torch.float32)
    out = {}
    old = hx.tune_get(_lib.KNOB_GEMM_F32)
    try:
        for v in (13, 14, 16, 17):
            hx.tune_set(_lib.KNOB_GEMM_F32, v)
            Cd = to_dev(torch, hx, np.pad(C, ((0, ld_pad), (0, 0))), torch.float32)
            ctx = hx.get_context()
            ctx.check(ctx.lib.hfpx_gemm_sub(ctx.handle, _lib.HFPX_F32, hx._ptr(Cd), ldc, hx._ptr(Ad), lda,
                                             hx._ptr(Bd), ldb, M, N, K))
            ctx.stream.synchronize()
            out[v] = Cd.cpu().numpy()[:M]
    finally:
        hx.tune_set(_lib.KNOB_GEMM_F32, old)
    assert np.array_equal(out[16], out[14])
    assert np.array_equal(out[17], out[13])


@pytest.mark.parametrize("scaled", [False, True])
def test_sgemm_bf16_split_is_fp32_accurate(env, scaled):
    """The three-way bf16 split product (variants 13, 14) against float64 on the
    same FP32 inputs: its error stays within a small factor of the FP32
    MFMA kernel's (variant 2), i.e. FP32-level, also when the rows of A and
    the columns of B span 10^-8 .. 10^8 (products stay inside FP32's range;
    a dropped low part would cost a factor 2^6 or more)."""
    torch, hx = env
    from hfpx3d_amd import _lib
    print(f"case: sgemm bf16 split accuracy scaled={scaled}", flush=True)
    M, N, K = 700, 650, 1024
    rng = np.random.default_rng(11)
    A, B, C = rng.normal(size=(M, K)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    if scaled:
        A *= 10.0 ** rng.uniform(-8, 8, size=(M, 1))
        B *= 10.0 ** rng.uniform(-8, 8, size=(1, N))
        C *= np.abs(A).max(axis=1, keepdims=True) * np.abs(B).max(axis=0, keepdims=True)
    A32, B32, C32 = (x.astype(np.float32).astype(np.float64) for x in (A, B, C))
    ref = C32 - A32 @ B32
    scale = (np.abs(A32) @ np.abs(B32)) + np.abs(C32)  # elementwise error scale
    err = {}
    old = hx.tune_get(_lib.KNOB_GEMM_F32)
    try:
        for v in (2, 13, 14, 15, 16, 17):
            hx.tune_set(_lib.KNOB_GEMM_F32, v)
            Cd = to_dev(torch, hx, C32, torch.float32)
            hx.gemm_sub(Cd, to_dev(torch, hx, A32, torch.float32), to_dev(torch, hx, B32, torch.float32))
            err[v] = (np.abs(Cd.cpu().numpy() - ref) / scale).max()
    finally:
        hx.tune_set(_lib.KNOB_GEMM_F32, old)
    assert err[2] < 1e-6
    assert err[13] < 3 * err[2] + 1e-7, err
    assert err[14] < 3 * err[2] + 1e-7, err
    assert err[15] < 3 * err[2] + 1e-7, err
    assert err[16] < 3 * err[2] + 1e-7, err
    assert err[17] < 3 * err[2] + 1e-7, err


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("M,N,K", [(1000, 900, 300), (384, 256, 1024), (2304, 1280, 77)])
def test_dgemm_variants(env, variant, M, N, K):
    """FP64 Schur-complement products through every HFPX_KNOB_GEMM_F64
    variant (interior tiles unguarded, edge bands guarded, K tails) against
    numpy, 1e-13 of the largest entry of C."""
    from hfpx3d_amd import _lib
    torch, hx = env
    print(f"case: dgemm variant {variant} {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M + N + K + variant)
    A, B, C = rng.normal(size=(M, K)), rng.normal(size=(K, N)), rng.normal(size=(M, N))
    ctx = hx.get_context()
    hx.tune_set(_lib.KNOB_GEMM_F64, variant, ctx)
    try:
        Cd = to_dev(torch, hx, C, torch.float64)
        hx.gemm_sub(Cd, to_dev(torch, hx, A, torch.float64), to_dev(torch, hx, B, torch.float64), ctx)
        got = Cd.cpu().numpy()
    finally:
        hx.tune_set(_lib.KNOB_GEMM_F64, 0, ctx)
    ref = C - A @ B
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("dt,n,nb", [("f32", 3000, 1024), ("f64", 3000, 1024), ("f32", 4500, 1024),
                                     ("f64", 1500, 256), ("f32", 2000, 256)])
def test_getrf_lookahead_is_bitwise_the_plain_loop(env, mode, dt, n, nb):
    """The look-ahead factorisation (next block factorised on the main
    stream while the rest of the trailing update runs on a second stream)
    against the plain blocked loop on the same matrix: identical pivots and
    bitwise identical factors; and the pivots are LAPACK's for FP64."""
    from hfpx3d_amd import _lib
    torch, hx = env
    print(f"case: look-ahead mode {mode} {dt} n={n} nb={nb}", flush=True)
    A = np.random.default_rng(n + nb).normal(size=(n, n))
    dtype = torch.float64 if dt == "f64" else torch.float32
    ctx = hx.get_context()
    out = {}
    hx.tune_set(_lib.KNOB_LU_NB, nb, ctx)
    try:
        for la in (1, mode):
            hx.tune_set(_lib.KNOB_LU_LOOKAHEAD, la, ctx)
            lu = hx.LU(to_dev(torch, hx, A, dtype))
            out[la] = (lu.A.cpu().numpy(), lu.ipiv.cpu().numpy())
    finally:
        hx.tune_set(_lib.KNOB_LU_LOOKAHEAD, 0, ctx)
        hx.tune_set(_lib.KNOB_LU_NB, 0, ctx)
    assert np.array_equal(out[1][1], out[mode][1])
    assert np.array_equal(out[1][0], out[mode][0])
    if dt == "f64":
        sl = pytest.importorskip("scipy.linalg")
        _, piv = sl.lu_factor(A)
        assert np.array_equal(out[mode][1], piv + 1)


def test_tune_rejects_unknown_variant(env):
    torch, hx = env
    from hfpx3d_amd import _lib
    with pytest.raises(_lib.HfpxError):
        hx.tune_set(_lib.KNOB_GEMM_F32, _lib.GEMM_F32_VARIANTS)
    with pytest.raises(_lib.HfpxError):
        hx.tune_set(99, 0)
    with pytest.raises(_lib.HfpxError):
        hx.tune_set(_lib.KNOB_LU_NB, 300)


@pytest.mark.parametrize("n", [1, 7, 64, 300, 1000, 2500])
def test_getrf_f64_matches_lapack(env, n):
    torch, hx = env
    sl = pytest.importorskip("scipy.linalg")
    print(f"case: getrf f64 n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    Ad = to_dev(torch, hx, A, torch.float64)
    lu = hx.LU(Ad)
    LU = lu.A.cpu().numpy()
    piv = lu.ipiv.cpu().numpy()
    lu2, piv2 = sl.lu_factor(A)
    assert lu.info == 0
    # identical partial pivoting (no near-ties in a random matrix)
    assert np.array_equal(piv - 1, piv2)
    assert np.abs(LU - lu2).max() < 1e-12 * max(1.0, np.abs(lu2).max()) * max(1, n / 100)
    b = rng.normal(size=(n, 3))
    x = lu.solve(to_dev(torch, hx, b, torch.float64)).cpu().numpy()
    assert np.abs(A @ x - b).max() / (np.abs(A).max() * np.abs(x).max() * n) < 1e-15 * 10


@pytest.mark.parametrize("n,nrhs", [(1, 1), (127, 3), (128, 64), (129, 65), (1000, 64), (2049, 130), (1500, 33)])
def test_getrs_f32_block_sweeps(env, n, nrhs):
    """FP32 solve with the GPU's own FP32 factors: the block-inverse sweeps
    against the same factors applied by LAPACK in float64 (lu_solve), within
    FP32 rounding times the conditioning of the triangular factors."""
    torch, hx = env
    sl = pytest.importorskip("scipy.linalg")
    print(f"case: getrs f32 n={n} nrhs={nrhs}", flush=True)
    rng = np.random.default_rng(5 * n + nrhs)
    A = rng.normal(size=(n, n)) + np.sqrt(n) * np.eye(n)
    lu = hx.LU(to_dev(torch, hx, A, torch.float32))
    LU = lu.A.cpu().numpy().astype(np.float64)
    piv = lu.ipiv.cpu().numpy() - 1
    b = rng.normal(size=(n, nrhs))
    x = lu.solve(to_dev(torch, hx, b, torch.float32)).cpu().numpy().astype(np.float64)
    xr = sl.lu_solve((LU, piv), b.astype(np.float32).astype(np.float64))
    assert np.abs(x - xr).max() / np.abs(xr).max() < 2e-5


@pytest.mark.parametrize("nb", [256, 512, 2048, 4096])
def test_getrf_outer_block_sizes_match_lapack(env, nb):
    """Every outer block size of the recursive LU gives LAPACK's pivots and
    factors (n = 2,500 spans several outer blocks for the small sizes and a
    partial one for the large)."""
    torch, hx = env
    sl = pytest.importorskip("scipy.linalg")
    from hfpx3d_amd import _lib
    print(f"case: getrf f64 NB={nb}", flush=True)
    n = 2500
    rng = np.random.default_rng(nb)
    A = rng.normal(size=(n, n))
    old = hx.tune_get(_lib.KNOB_LU_NB)
    hx.tune_set(_lib.KNOB_LU_NB, nb)
    try:
        lu = hx.LU(to_dev(torch, hx, A, torch.float64))
    finally:
        hx.tune_set(_lib.KNOB_LU_NB, old)
    lu2, piv2 = sl.lu_factor(A)
    assert np.array_equal(lu.ipiv.cpu().numpy() - 1, piv2)
    assert np.abs(lu.A.cpu().numpy() - lu2).max() < 1e-12 * np.abs(lu2).max() * n / 100


@pytest.mark.parametrize("dt,n", [("f64", 2500), ("f32", 2500), ("f32", 5000), ("f64", 1100)])
def test_getrf_panel_kernels_agree(env, dt, n):
    """The exchange panel (tagged words, no grid barrier) and the grid-
    barrier panel do the same arithmetic in the same order: bitwise equal
    factors and pivots, and LAPACK's pivots in FP64.  n = 5,000 takes five
    panel workgroups, n = 1,100 two (the second mostly idle)."""
    torch, hx = env
    sl = pytest.importorskip("scipy.linalg")
    from hfpx3d_amd import _lib
    print(f"case: getrf panels {dt} n={n}", flush=True)
    dtype = torch.float64 if dt == "f64" else torch.float32
    A = np.random.default_rng(n + 17).normal(size=(n, n))
    out = {}
    try:
        for knob in (1, 2):
            hx.tune_set(_lib.KNOB_PANEL, knob)
            lu = hx.LU(to_dev(torch, hx, A, dtype))
            out[knob] = (lu.A.cpu().numpy(), lu.ipiv.cpu().numpy())
    finally:
        hx.tune_set(_lib.KNOB_PANEL, 0)
    assert np.array_equal(out[1][1], out[2][1])
    assert np.array_equal(out[1][0], out[2][0])
    if dt == "f64":
        lu2, piv2 = sl.lu_factor(A)
        assert np.array_equal(out[2][1] - 1, piv2)


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("panel", [1, 2])
def test_getrf_pivot_ties_take_first_row(env, dt, panel):
    """Equal |a| in the first column, spread over several panel workgroups
    (n = 3,000: rows 1,500, 1,700 (negative), 2,100 and 2,999): the first
    such row is the pivot, as in LAPACK (idamax takes the first maximum)."""
    torch, hx = env
    from hfpx3d_amd import _lib
    print(f"case: getrf ties {dt} panel={panel}", flush=True)
    n = 3000
    A = np.random.default_rng(3).uniform(-1, 1, size=(n, n))
    for r, v in ((2100, 5.0), (1700, -5.0), (2999, 5.0), (1500, 5.0)):
        A[r, 0] = v
    try:
        hx.tune_set(_lib.KNOB_PANEL, panel)
        lu = hx.LU(to_dev(torch, hx, A, torch.float64 if dt == "f64" else torch.float32))
    finally:
        hx.tune_set(_lib.KNOB_PANEL, 0)
    assert int(lu.ipiv.cpu().numpy()[0]) == 1501


@pytest.mark.parametrize("n", [300, 2049, 5000])
def test_getrf_f32_backward_error(env, n):
    """FP32 factors of a random matrix: |PA - LU| within n FP32 roundings
    and |L| <= 1."""
    torch, hx = env
    print(f"case: getrf f32 n={n}", flush=True)
    rng = np.random.default_rng(n + 1)
    A = rng.normal(size=(n, n)).astype(np.float32)
    lu = hx.LU(to_dev(torch, hx, A, torch.float32))
    LU = lu.A.cpu().numpy().astype(np.float64)
    piv = lu.ipiv.cpu().numpy() - 1
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.astype(np.float64).copy()
    for j, p in enumerate(piv):
        PA[[j, p]] = PA[[p, j]]
    assert np.abs(PA - L @ U).max() / np.abs(A).max() < n * 1.2e-7
    assert np.abs(L).max() <= 1.0 + 1e-6  # partial pivoting bound


def test_getrf_reports_singular(env):
    torch, hx = env
    print("case: singular", flush=True)
    # an exactly zero column stays exactly zero under elimination, so the
    # first zero pivot is that column (a duplicated column is not exactly
    # singular after rounded elimination, in LAPACK either)
    A = np.random.default_rng(3).normal(size=(50, 50))
    A[:, 17] = 0.0
    st = hx.Status()
    lu = hx.LU(to_dev(torch, hx, A, torch.float64), status=st)
    assert not st.ok() and st.error == "MatrixSingular"
    o = oracle_lib.oracle()
    _, _, info = o.getrf(A)
    # same first zero pivot as the oracle's LAPACK-semantics LU
    assert lu.info == info == 18
    assert st.info == 17


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_getrf_reports_panel_timeout_promptly(env, dt):
    """The panel's bounded waits, forced from the host: with the time-out
    flag raised before the factorisation (HFPX_KNOB_TEST_PANEL_TIMEOUT)
    every panel returns at entry, every kernel exits normally, and getrf
    reports HFPX_ERR_HIP at once instead of after the waits' bounds.  The
    next factorisation on the same context is correct again."""
    import time
    from hfpx3d_amd import _lib
    torch, hx = env
    print(f"case: panel time-out path {dt}", flush=True)
    n = 3000
    A = np.random.default_rng(9).normal(size=(n, n))
    dtype = torch.float64 if dt == "f64" else torch.float32
    ctx = hx.get_context()
    hx.tune_set(_lib.KNOB_TEST_PANEL_TIMEOUT, 1, ctx)
    try:
        t0 = time.perf_counter()
        with pytest.raises(_lib.HfpxError) as e:
            hx.LU(to_dev(torch, hx, A, dtype))
        dt_s = time.perf_counter() - t0
    finally:
        hx.tune_set(_lib.KNOB_TEST_PANEL_TIMEOUT, 0, ctx)
    assert e.value.status == _lib.HFPX_ERR_HIP
    assert "timed out" in str(e.value)
    assert dt_s < 10.0, dt_s
    lu = hx.LU(to_dev(torch, hx, A, dtype))
    sl = pytest.importorskip("scipy.linalg")
    _, piv = sl.lu_factor(A.astype(np.float32) if dt == "f32" else A)
    if dt == "f64":
        assert np.array_equal(lu.ipiv.cpu().numpy(), piv + 1)


def test_bem_penny121_solution_matches_oracle(env):
    """Assembly + LU + solve on the 121-element penny (tip_type 0, the C1
    configuration E = 1, nu = 0.3, p = 1): DD agree with the oracle's FP64
    solve and with Sneddon's opening away from the tip."""
    torch, hx = env
    print("case: penny121 solve", flush=True)
    nods = np.load(golden("Nodes_pennymesh121el_32.npy"))
    conn = np.load(golden("Elems_pennymesh121el_32.npy")).astype(np.int64)
    conn[:3] -= 1
    G, nu = 1 / 2.6, 0.3
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    npar = hx.Num_Param_T(beta=0.125, tip_type=0)
    md = hx.Mesh_Data_T(mesh=mesh)
    sae = hx.SAE_T()
    sae.matrix = hx.make_3dbem_matrix_s(G, nu, mesh, npar, md.dof_h_dd)
    hx.add_s_inf_to_3dbem_rhs(md, hx.Load_T(s_inf=np.array([1.0, 1, 1, 0, 0, 0])), sae)
    st = hx.Status()
    lu = hx.LU(sae.matrix, status=st)
    st.abortOnError()
    x = lu.solve(sae.rhs_v).cpu().numpy()
    o = oracle_lib.oracle()
    Ao, dof = o.assemble(nods, conn, G, nu)
    r = o.rhs(nods, conn, dof, np.array([1.0, 1, 1, 0, 0, 0]))
    LUo, piv, info = o.getrf(Ao)
    xo = o.getrs(LUo, piv, r)
    assert np.abs(x - xo).max() / np.abs(xo).max() < 1e-12
    dd = hx.write_dd_vector_to_md(x, md.dof_h_dd)
    # opening against Sneddon 4 (1 - nu) p sqrt(1 - r^2) / (pi G) away from the tip
    err = 0.0
    for e in range(conn.shape[1]):
        v = nods[:3, conn[:3, e]]
        _, _, _, cp = o.el_geom(v, 0.0)
        for k in range(6):
            rr = np.hypot(cp[0, k], cp[1, k])
            if rr > 0.8:
                continue
            s = o.sneddon_dd(G, nu, 1.0, 1.0, cp[:, k])
            err = max(err, abs(dd[6 * e + k, 2] - s[2]) / s[2])
    assert err < 0.03
