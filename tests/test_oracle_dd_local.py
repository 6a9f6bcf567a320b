"""Pins the oracle's is_dd_local = true branches (DD in each element's own
frame; system_assembly.cpp:186-195, 358-387, 529-536).  CPU only.

Pins: the reference build with is_dd_local = true (assembly, stress,
volume-control border); the identities A_local = A_global . blockdiag(R_e^T)
and stress(local DD R d) = stress(global DD d)."""
import numpy as np
import pytest

import oracle_lib
from oracle_lib import golden
from test_oracle import assert_blocks_match_or_reference_wrong, load_mesh, rel

O = oracle_lib.oracle()
REF = oracle_lib.reference_or_none()
needs_ref = pytest.mark.skipif(REF is None, reason="reference build oracle/_ref not present")


def frames(nods, conn):
    return [O.el_geom(nods[:3, conn[:3, e]])[1] for e in range(conn.shape[1])]


def meshes():
    d = np.load(golden("ref_assembly.npz"))
    nods, conn = load_mesh("pennymesh24el_32")
    return {"penny24": (nods, conn, 1.0, 0.35), "wavy24": (d["wavy24_nods"], d["wavy24_conn"], 1 / 2.6, 0.3)}


@pytest.mark.parametrize("case", ["penny24", "wavy24"])
def test_assembly_dd_local_is_global_times_frames(case):
    nods, conn, G, nu = meshes()[case]
    Ag, dof = O.assemble(nods, conn, G, nu, tip_type=0)
    Al, _ = O.assemble(nods, conn, G, nu, tip_type=0, dd_local=True)
    R = frames(nods, conn)
    want = Ag.copy()
    for e in range(conn.shape[1]):
        for n in range(6):
            c = slice(18 * e + 3 * n, 18 * e + 3 * n + 3)
            want[:, c] = Ag[:, c] @ R[e].T
    assert rel(Al, want) < 1e-14
    assert rel(Al, Ag) > 1e-3      # the branch changes the matrix


@needs_ref
@pytest.mark.parametrize("case,tip", [("penny24", 1), ("wavy24", 0)])
def test_assembly_dd_local_matches_reference_build(case, tip):
    nods, conn, G, nu = meshes()[case]
    A, dof = O.assemble(nods, conn, G, nu, tip_type=tip, dd_local=True)
    Ar, dof_r = REF.assemble(nods, conn, G, nu, tip_type=tip, dd_local=True)
    assert np.array_equal(dof, dof_r)
    if tip == 0:
        assert_blocks_match_or_reference_wrong(A, Ar, nods, conn, G, nu)
    else:
        assert rel(A, Ar) < 1e-9


def test_stress_dd_local_equals_rotated_global():
    nods, conn = load_mesh("pennymesh24el_32")
    d = np.load(golden("ref_rhs_stress.npz"))
    R = frames(nods, conn)
    ddg = d["dd"]
    ddl = ddg.copy()
    for e in range(conn.shape[1]):
        ddl[6 * e:6 * e + 6] = ddg[6 * e:6 * e + 6] @ R[e].T
    Sg = O.stress(nods, conn, 1.0, 0.35, ddg, d["pts"], d["s_inf"])
    Sl = O.stress(nods, conn, 1.0, 0.35, ddl, d["pts"], d["s_inf"], dd_local=True)
    assert np.abs(Sl - Sg).max() <= 1e-12 * np.abs(Sg).max()


@needs_ref
def test_stress_dd_local_matches_reference_build():
    nods, conn = load_mesh("pennymesh24el_32")
    d = np.load(golden("ref_rhs_stress.npz"))
    S = O.stress(nods, conn, 1.0, 0.35, d["dd"], d["pts"], d["s_inf"], dd_local=True)
    Sr = REF.stress(nods, conn, 1.0, 0.35, d["dd"], d["pts"], d["s_inf"], dd_local=True)
    e = np.abs(S - Sr).max(axis=1) / np.abs(Sr).max()
    assert np.median(e) < 1e-13
    assert e.max() < 1e-9


@needs_ref
@pytest.mark.parametrize("tip", [0, 1])
def test_vc_border_dd_local_matches_reference_build(tip):
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof = REF.matrix_vc(nods, conn, 1.0, 0.35, tip_type=tip, dd_local=True)
    n = Avc.shape[0] - 1
    row, col = O.vc_border(nods, conn, dof, dd_local=True)
    assert np.abs(Avc[n, :n] - row).max() <= 1e-13 * np.abs(row).max()
    assert np.array_equal(Avc[:n, n], col)
    # only opening components carry volume
    comp = np.zeros(n, dtype=int)
    for e in range(conn.shape[1]):
        for k in range(18):
            if dof[e, k] >= 0:
                comp[dof[e, k]] = k % 3
    # (vertex entries are cancellation noise around 0)
    assert (row[comp != 2] == 0).all() and (row[comp == 2] >= -1e-13 * row.max()).all()
