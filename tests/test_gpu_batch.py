"""Batches of independent cracks (no reference counterpart): every crack's
matrix and right-hand side against the single-crack path, every solution
against LAPACK on that crack's FP64 matrix (FP64 group and FP32 + refinement
group), and hipGraph replays against the eager launch sequence."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d
    return torch, hfp3d


def make_cracks(n_cracks, seed, n_min, n_max):
    """Planar elliptical cracks of uneven size and random orientation, E = 1,
    nu and pressure p drawn per crack (load s_inf = p I)."""
    from hfpx3d_amd import meshgen
    out = []
    for c in meshgen.batch_cracks(n_cracks, seed, n_min, n_max):
        p = c["p"]
        out.append(dict(nods=c["nods"], conn=c["conn"], G=1.0 / (2.0 * (1.0 + c["nu"])), nu=c["nu"],
                        s_inf=np.array([p, p, p, 0.0, 0.0, 0.0])))
    return out


def single_crack(hx, c):
    """The crack through the single-crack path: FP64 matrix and rhs."""
    mesh = hx.Mesh_Geom_T(nods=c["nods"], conn=c["conn"])
    md = hx.Mesh_Data_T(mesh=mesh)
    npar = hx.Num_Param_T(beta=0.125, tip_type=0)
    sae = hx.SAE_T()
    sae.matrix = hx.make_3dbem_matrix_s(c["G"], c["nu"], mesh, npar, md.dof_h_dd)
    hx.add_s_inf_to_3dbem_rhs(md, hx.Load_T(s_inf=c["s_inf"]), sae)
    return sae.matrix.cpu().numpy(), sae.rhs_v.cpu().numpy()


def test_batch_assembly_matches_single_crack(env):
    torch, hx = env
    print("case: batch assembly 10 cracks", flush=True)
    cracks = make_cracks(10, 7, 20, 90)
    bt = hx.Batch(cracks, mixed_min_n=900)
    bt.assemble()
    for i, c in enumerate(cracks):
        A1, b1 = single_crack(hx, c)
        A = bt.matrix(i)
        assert A.shape == A1.shape
        assert np.abs(A - A1).max() <= 1e-15 * np.abs(A1).max(), i
        assert np.abs(bt.rhs(i) - b1).max() <= 1e-15 * np.abs(b1).max(), i


@pytest.mark.parametrize("use_graph,bgemm", [(False, 0), (True, 0), (False, 1), (False, 2)])
def test_batch_solve_matches_lapack(env, use_graph, bgemm):
    """n from 360 to 1,620 DoF; cracks with n >= 900 take FP32 LU + 3
    refinement steps, the others FP64 LU.  bgemm selects the FP32 updates
    (0 default: outer updates as bf16-split products; 1: FP32 MFMA with a
    read-modify-write epilogue; 2: FP32 MFMA with C preloaded)."""
    torch, hx = env
    from hfpx3d_amd import _lib
    sl = pytest.importorskip("scipy.linalg")
    print(f"case: batch solve graph={use_graph} bgemm={bgemm}", flush=True)
    cracks = make_cracks(10, 7, 20, 90)
    bt = hx.Batch(cracks, mixed_min_n=900)
    hx.tune_set(_lib.KNOB_BGEMM_F32, bgemm)
    try:
        bt.solve(n_refine=3, use_graph=use_graph)
    finally:
        hx.tune_set(_lib.KNOB_BGEMM_F32, 0)
    groups = set()
    for i, c in enumerate(cracks):
        A1, b1 = single_crack(hx, c)
        xr = sl.solve(A1, b1, check_finite=False)
        x, berr, info = bt.solution(i)
        mixed = A1.shape[0] >= 900
        groups.add(mixed)
        assert info == 0
        assert berr < 1e-14, (i, berr)
        assert np.abs(x - xr).max() / np.abs(xr).max() < (1e-10 if mixed else 1e-12), i
        # solving leaves the assembled FP64 matrix in place (the FP64 group
        # is factorised in a working copy)
        assert np.abs(bt.matrix(i) - A1).max() <= 1e-15 * np.abs(A1).max()
    assert groups == {False, True}


def test_batch_graph_replays_match_eager(env):
    """Three graph replays, with other GPU work and synchronisations in
    between, give bitwise the eager result (all kernels are deterministic)."""
    torch, hx = env
    print("case: batch graph replays", flush=True)
    cracks = make_cracks(6, 11, 20, 70)
    bt = hx.Batch(cracks, mixed_min_n=700)
    bt.solve(n_refine=2, use_graph=False)
    ref = [bt.solution(i)[0] for i in range(len(cracks))]
    for rep in range(3):
        bt.solve(n_refine=2, use_graph=True)
        y = torch.randn(1000, 1000, device="cuda") @ torch.randn(1000, 1000, device="cuda")
        torch.cuda.synchronize()
        del y
        for i in range(len(cracks)):
            assert np.array_equal(bt.solution(i)[0], ref[i]), (rep, i)
