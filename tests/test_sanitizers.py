"""The CPU-side code under AddressSanitizer + UndefinedBehaviorSanitizer:
the C++ host layer (config reading, npy / CSV mesh loaders including corrupt
files, DoF handle) and the C oracle (every routine once), built from
tests/native/Makefile with g++ / gcc.  No GPU code is instrumented; a
sanitizer report fails the run."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import golden
from test_host import REF_CONFIG

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")


@pytest.fixture(scope="module")
def asan_build():
    r = subprocess.run(["make", "-C", NATIVE, "-s", "asan"], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        pytest.fail(f"sanitizer build failed:\n{r.stdout}\n{r.stderr}")
    return os.path.join(NATIVE, "_asan")


def run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, env=ENV)
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode not in (23, 24), r.stderr[-4000:]
    return r


def test_oracle_under_sanitizers(asan_build):
    r = run(os.path.join(asan_build, "oracle_check"))
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout)
    assert out["n_dof"] == 108 and out["info"] == 0 and out["vc"] == 0


def test_host_config_and_loaders_under_sanitizers(asan_build, tmp_path):
    exe = os.path.join(asan_build, "host_check")
    mesh = tmp_path / "mesh"
    mesh.mkdir()
    for kind in ("Nodes", "Elems"):
        np.save(mesh / f"{kind}_pennymesh24el_32.npy", np.load(golden(f"{kind}_pennymesh24el_32.npy")))
    cfg = tmp_path / "config.toml"
    cfg.write_text(REF_CONFIG.format(conn="Elems_pennymesh24el_32.npy", nods="Nodes_pennymesh24el_32.npy",
                                     fmt="npy32", fix_tip=1))
    r = run(exe, str(cfg))
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    out = json.loads(r.stdout)
    assert out["conn"] == [5, 24] and out["n_dof"] > 0


def _npy(path, header, payload):
    h = header.encode()
    pad = 64 - (10 + len(h) + 1) % 64
    path.write_bytes(b"\x93NUMPY\x01\x00" + (len(h) + pad + 1).to_bytes(2, "little") + h + b" " * pad + b"\n" +
                     payload)


CORRUPT = [
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, 99999999999), }", b"\0" * 72),
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, -5), }", b"\0" * 72),
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, x), }", b"\0" * 72),
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, 3", b"\0" * 72),
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (3, 4), }", b"\0" * 40),
    ("{'descr': '<c16', 'fortran_order': False, 'shape': (3, 3), }", b"\0" * 144),
    ("{'descr': '<f8', 'fortran_order': False, 'shape': (9999999999999, 3), }", b"\0" * 72),
    ("{'descr'", b""),
]


@pytest.mark.parametrize("k", range(len(CORRUPT)))
def test_npy_loader_rejects_corrupt_files_under_sanitizers(asan_build, tmp_path, k):
    """Malformed headers, impossible shapes and truncated data come back as
    loader errors, with no out-of-bounds access or exception escaping."""
    bad = tmp_path / "bad.npy"
    _npy(bad, *CORRUPT[k])
    r = run(os.path.join(asan_build, "host_check"), "--npy-mesh", str(bad), golden("Elems_pennymesh24el_32.npy"))
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr[-2000:])
    assert "error" in json.loads(r.stdout)


def test_npy_loader_reads_reference_meshes_under_sanitizers(asan_build):
    r = run(os.path.join(asan_build, "host_check"), "--npy-mesh", golden("Nodes_pennymesh121el_32.npy"),
            golden("Elems_pennymesh121el_32.npy"))
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout)
    assert out["conn"][1] == 121 and out["nods"][1] == 75
