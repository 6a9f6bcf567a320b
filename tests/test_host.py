"""C++ host layer (config.toml semantics, mesh loaders, DoF handle) on the
CPU, through tests/native/hfpx3d_host_check (no device call)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib
from oracle_lib import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "native", "hfpx3d_host_check")

# the reference's own configuration keys (config.toml), including the ones
# its reader ignores (injection / fluid / contact)
REF_CONFIG = """
mesh_input_directory = "mesh"
mesh_conn_fname = "{conn}"
nodes_crd_fname = "{nods}"
output_signature = "_psc_24_ele"
mesh_input_format = "{fmt}"
do_save_matrix = false
do_save_solution = true
do_postprocess = true
output_directory = "out"
output_format = "csv"
observ_crd_directory = "obs"
observ_crd_fname = "observation_points.csv"
observ_input_format = "csv"
array_origin = 1
cp_offset = 0.125
fix_tip = {fix_tip}
is_dd_local = false
is_tr_local = false
S_xx = 0
S_yy = 0
S_zz = -1.0
S_xy = 0
S_xz = 0.25
S_yz = 0
Ambient_pore_pressure = 0
Constant_overpressure = 3.0e-1
number_solids = 1
Shear_modulus = 1.0
Poisson_ratio = 3.5e-1
density = 1.0               # density is in g*cm^(-3)
Contact_model_id = 0 # Dugdale cohesion / Mohr-Coulomb friction
"""


def run_check(cfg_path):
    r = subprocess.run([CHECK, cfg_path], capture_output=True, text=True, timeout=60)
    return json.loads(r.stdout)


def write_case(tmp_path, fmt="npy32", fix_tip=1, extra="", body=None):
    mesh_dir = tmp_path / "mesh"
    mesh_dir.mkdir(exist_ok=True)
    nods = np.load(golden("Nodes_pennymesh24el_32.npy"))
    conn = np.load(golden("Elems_pennymesh24el_32.npy"))  # 1-based
    if fmt == "npy32":
        cn, nn = "Elems.npy", "Nodes.npy"
        np.save(mesh_dir / cn, np.asfortranarray(conn.astype(np.int32)))
        np.save(mesh_dir / nn, np.asfortranarray(nods))
    elif fmt == "npy64":
        cn, nn = "Elems64.npy", "Nodes64.npy"
        np.save(mesh_dir / cn, np.asfortranarray(conn.astype(np.int64)))
        np.save(mesh_dir / nn, np.ascontiguousarray(nods))  # C order also accepted
    else:
        cn, nn = "Elems.csv", "Nodes.csv"
        np.savetxt(mesh_dir / cn, conn.T, fmt="%d", delimiter=",")
        np.savetxt(mesh_dir / nn, nods.T, fmt="%.17g", delimiter=",")
    text = body if body is not None else REF_CONFIG.format(conn=cn, nods=nn, fmt=fmt, fix_tip=fix_tip)
    (tmp_path / "config.toml").write_text(text + extra)
    return str(tmp_path / "config.toml"), nods, conn


def mesh_sums(nods, conn0):
    ns = sum(nods[i, j] * (1 + i + 3 * j) for j in range(nods.shape[1]) for i in range(nods.shape[0]))
    cs = sum(float(conn0[i, j]) * (1 + i + 5 * j) for j in range(conn0.shape[1]) for i in range(conn0.shape[0]))
    return ns, cs


@pytest.mark.parametrize("fmt", ["npy32", "npy64", "csv"])
def test_reference_config_semantics(tmp_path, fmt):
    cfg, nods, conn = write_case(tmp_path, fmt)
    d = run_check(cfg)
    assert "error" not in d, d
    assert d["G"] == 1.0 and d["nu"] == 0.35 and d["beta"] == 0.125 and d["tip_type"] == 1
    # normal components flipped (compression positive in the file), shear kept,
    # integer literals for float keys ignored (S_xx = 0 stays at its default)
    assert d["s_inf"] == [0.0, 0.0, 1.0, 0.0, 0.25, 0.0]
    assert d["save_solution"] == 1 and d["postprocess"] == 1 and d["save_matrix"] == 0
    conn0 = conn.astype(np.int64).copy()
    conn0[:3] -= 1  # array_origin = 1 shifts the vertex rows only
    ns, cs = mesh_sums(nods, conn0)
    assert d["nods"] == list(nods.shape) and d["conn"] == list(conn.shape)
    assert abs(d["nods_sum"] - ns) <= 1e-9 * abs(ns) and d["conn_sum"] == cs
    n, dof = oracle_lib.oracle().dof_h(nods, conn0, 1)
    assert d["n_dof"] == n == 342
    assert d["dof_sum"] == float(sum(dof[e, c] * (1 + c + 18 * e) for e in range(dof.shape[0]) for c in range(18)))
    assert d["input_dir"].endswith("/mesh/") and d["obs_dir"].endswith("/obs/") and d["out_dir"].endswith("/out/")
    assert d["out_file"] == "test_solution_psc_24_ele.csv" and d["strs_file"] == "test_stresses_psc_24_ele.csv"


def test_solid_table_young_modulus(tmp_path):
    body = REF_CONFIG.format(conn="Elems.npy", nods="Nodes.npy", fmt="npy32", fix_tip=0)
    body = body.replace("Shear_modulus = 1.0\n", "").replace("Poisson_ratio = 3.5e-1\n", "")
    body += "\n[solid]\nYoung_modulus = 1.0\nPoisson_ratio = 0.3\n"
    cfg, _, _ = write_case(tmp_path, body=body)
    d = run_check(cfg)
    assert abs(d["G"] - 1 / 2.6) < 1e-15 and d["nu"] == 0.3 and d["tip_type"] == 0 and d["n_dof"] == 432


def test_integer_literal_for_float_key_is_ignored(tmp_path):
    body = REF_CONFIG.format(conn="Elems.npy", nods="Nodes.npy", fmt="npy32", fix_tip=1)
    body = body.replace("Shear_modulus = 1.0", "Shear_modulus = 2").replace("cp_offset = 0.125", "cp_offset = 0")
    cfg, _, _ = write_case(tmp_path, body=body)
    d = run_check(cfg)
    assert d["G"] == 1.0 and d["beta"] == 0.125


def test_missing_mesh_file_is_reported_not_aborted(tmp_path):
    body = REF_CONFIG.format(conn="nope.npy", nods="Nodes.npy", fmt="npy32", fix_tip=1)
    cfg, _, _ = write_case(tmp_path, body=body)
    r = subprocess.run([CHECK, cfg], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in json.loads(r.stdout)["error"]


def test_missing_required_key_is_reported(tmp_path):
    body = REF_CONFIG.format(conn="Elems.npy", nods="Nodes.npy", fmt="npy32", fix_tip=1)
    body = body.replace('mesh_conn_fname = "Elems.npy"\n', "")
    cfg, _, _ = write_case(tmp_path, body=body)
    r = subprocess.run([CHECK, cfg], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "mesh_conn_fname" in r.stdout


def test_save_data_to_npy(tmp_path):
    """save_data_to_npy writes .npy v1.0 files numpy reads back: FP64 exactly,
    FP32 as the rounding, column-major with the array's shape."""
    r = subprocess.run([CHECK, "--npy", str(tmp_path) + "/"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    want = np.array([[i + 10.0 * j + 0.1 for j in range(5)] for i in range(3)])
    a = np.load(tmp_path / "check64.npy", allow_pickle=False)
    b = np.load(tmp_path / "check32.npy", allow_pickle=False)
    assert a.dtype == np.float64 and a.shape == (3, 5) and a.flags["F_CONTIGUOUS"]
    assert b.dtype == np.float32 and b.shape == (3, 5)
    assert np.array_equal(a, want) and np.array_equal(b, want.astype(np.float32))
    raw = (tmp_path / "check64.npy").read_bytes()
    assert raw[:8] == b"\x93NUMPY\x01\x00" and (10 + int.from_bytes(raw[8:10], "little")) % 64 == 0
