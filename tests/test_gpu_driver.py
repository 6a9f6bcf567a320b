"""The main()-equivalent driver (hfpx3d_amd/hfpx3d_main) end to end on the
GPU against the reference's own example outputs (Test_Output/): the 24- and
121-element penny cracks with the parameters of the reference config
(G = 1, nu = 0.35, S_zz = -1, fix_tip = 1, beta = 0.125), solution (nodes +
DD) and stresses at the 11 observation points, written at 6 significant
digits by the reference."""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import golden
from test_host import REF_CONFIG

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "hfpx3d_amd", "hfpx3d_main")


@pytest.mark.parametrize("stem,tag", [("pennymesh24el_32", "psc_24"), ("pennymesh121el_32", "psc_121")])
def test_driver_reproduces_reference_outputs(tmp_path, stem, tag):
    print(f"case: driver {tag}", flush=True)
    # an LU in this process first: the driver's own LU then runs while this
    # process holds the device (its cooperative panel launches included)
    import torch
    from hfpx3d_amd import hfp3d as hx
    ctx = hx.get_context()
    A, _ = hx.alloc_colmajor(300, 300, torch.float64, ctx)
    with torch.cuda.stream(ctx.stream):
        A.copy_(torch.randn(300, 300, dtype=torch.float64) + 20 * torch.eye(300, dtype=torch.float64))
    hx.LU(A)
    torch.cuda.synchronize()
    (tmp_path / "mesh").mkdir()
    (tmp_path / "obs").mkdir()
    (tmp_path / "out").mkdir()
    np.save(tmp_path / "mesh" / "Elems.npy", np.load(golden(f"Elems_{stem}.npy")))
    np.save(tmp_path / "mesh" / "Nodes.npy", np.load(golden(f"Nodes_{stem}.npy")))
    np.savetxt(tmp_path / "obs" / "observation_points.csv", np.load(golden("testoutput_observation_points.npy")),
               delimiter=",", fmt="%.17g")
    cfg = REF_CONFIG.format(conn="Elems.npy", nods="Nodes.npy", fmt="npy32", fix_tip=1)
    cfg = cfg.replace("S_xz = 0.25", "S_xz = 0").replace("_psc_24_ele", f"_{tag}_ele")
    (tmp_path / "config.toml").write_text(cfg)
    r = subprocess.run([MAIN, str(tmp_path / "config.toml")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "Assembly:" in r.stdout and "Solution:" in r.stdout and "Fixed DoF" in r.stdout
    sol = np.loadtxt(tmp_path / "out" / f"test_solution_{tag}_ele.csv", delimiter=",")
    ref = np.load(golden(f"testoutput_test_solution_{tag}_ele.npy"))
    assert sol.shape == ref.shape
    assert np.abs(sol[:, :3] - ref[:, :3]).max() < 1e-5          # node coordinates
    assert np.abs(sol[:, 4:] - ref[:, 4:]).max() < 2e-6          # DD (6 digits)
    st = np.loadtxt(tmp_path / "out" / f"test_stresses_{tag}_ele.csv", delimiter=",")
    sref = np.load(golden(f"testoutput_test_stresses_{tag}_ele.npy"))
    assert np.abs(st - sref).max() < 2e-5 * np.abs(sref).max()


@pytest.mark.parametrize("fmt", ["npy64", "npy32"])
def test_driver_writes_npy_outputs(tmp_path, fmt):
    """output_format npy64 / npy32 (empty branches in the reference's main):
    the same solution, stresses and matrix as the CSV run, in .npy files."""
    print(f"case: driver npy output {fmt}", flush=True)
    import torch
    from hfpx3d_amd import hfp3d as hx
    ctx = hx.get_context()
    torch.cuda.synchronize()
    (tmp_path / "mesh").mkdir()
    (tmp_path / "obs").mkdir()
    (tmp_path / "out").mkdir()
    stem, tag = "pennymesh24el_32", "psc_24"
    np.save(tmp_path / "mesh" / "Elems.npy", np.load(golden(f"Elems_{stem}.npy")))
    np.save(tmp_path / "mesh" / "Nodes.npy", np.load(golden(f"Nodes_{stem}.npy")))
    np.savetxt(tmp_path / "obs" / "observation_points.csv", np.load(golden("testoutput_observation_points.npy")),
               delimiter=",", fmt="%.17g")
    cfg = REF_CONFIG.format(conn="Elems.npy", nods="Nodes.npy", fmt="npy32", fix_tip=1)
    cfg = cfg.replace("S_xz = 0.25", "S_xz = 0").replace('output_format = "csv"', f'output_format = "{fmt}"')
    cfg = cfg.replace("do_save_matrix = false", "do_save_matrix = true")
    # the stress points are read as CSV only when output_format is "csv"
    # (the reference selects the point loader by the output format)
    np.save(tmp_path / "obs" / "observation_points.npy", np.load(golden("testoutput_observation_points.npy")).T.copy())
    cfg = cfg.replace('observ_crd_fname = "observation_points.csv"', 'observ_crd_fname = "observation_points.npy"')
    (tmp_path / "config.toml").write_text(cfg)
    r = subprocess.run([MAIN, str(tmp_path / "config.toml")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    dt = np.float64 if fmt == "npy64" else np.float32
    sol = np.load(tmp_path / "out" / f"test_solution_{tag}_ele.npy", allow_pickle=False)
    assert sol.dtype == dt
    ref = np.load(golden(f"testoutput_test_solution_{tag}_ele.npy"))
    assert sol.shape == ref.shape
    assert np.abs(sol[:, 4:] - ref[:, 4:]).max() < 2e-6
    st = np.load(tmp_path / "out" / f"test_stresses_{tag}_ele.npy", allow_pickle=False)
    sref = np.load(golden(f"testoutput_test_stresses_{tag}_ele.npy"))
    assert st.dtype == dt and np.abs(st - sref).max() < 2e-5 * np.abs(sref).max()
    # the matrix file holds the driver's FP64 matrix (or its FP32 rounding)
    M = np.load(tmp_path / "out" / f"test_assembly_{tag}_ele.npy", allow_pickle=False)
    nods = np.load(golden(f"Nodes_{stem}.npy"))
    conn = np.load(golden(f"Elems_{stem}.npy")).astype(np.int64)
    conn[:3] -= 1
    A = hx.make_3dbem_matrix_s(1.0, 0.35, hx.Mesh_Geom_T(nods=nods, conn=conn),
                               hx.Num_Param_T(beta=0.125, tip_type=1), hx.DoF_Handle_T(), ctx=ctx).cpu().numpy()
    assert M.dtype == dt and M.shape == A.shape
    assert np.array_equal(M, A.astype(dt))
