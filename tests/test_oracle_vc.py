"""Pins the oracle's volume-control functions (hfpo_sf_integral,
hfpo_vc_border, hfpo_mod_system_vc) before they check the HIP path.  CPU only.

Pins used, per oracle function:
  sf_integral     closed form (int N_vertex dA = 0, int N_mid-edge dA = A/3),
                  reference build (oracle/_ref, el_p2_sf_integral)
  vc_border       reference build (make_3dbem_matrix_vc), volume of a uniform
                  unit opening = crack area (independent cross-product areas)
  mod_system_vc   reference build (mod_3dbem_system_vc), the selection form
                  M = P^T A P of the slot incidence P (numpy), the identity
                  truncation, merged DoF (sums), invalid inputs
"""
import numpy as np
import pytest

import oracle_lib
from test_oracle import load_mesh

O = oracle_lib.oracle()
REF = oracle_lib.reference_or_none()
needs_ref = pytest.mark.skipif(REF is None, reason="reference build oracle/_ref not present")


def tri_area(v):
    return 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]))


def crack_areas(nods, conn):
    return np.array([tri_area(nods[:3, conn[:3, e]]) for e in range(conn.shape[1])])


def incidence(dof_used, dof_orig, n_used, n_orig):
    """P[o, u] = number of slots (e, j) with used DoF u and original DoF o."""
    P = np.zeros((n_orig, n_used))
    for u, o in zip(dof_used.reshape(-1, order="F"), dof_orig.reshape(-1, order="F")):
        if u >= 0:
            P[o, u] += 1
    return P


# ---------------------------------------------------------------- sf integral

def test_sf_integral_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(100):
        v = rng.normal(size=(3, 3))
        sfm, rt, tau, cp = O.el_geom(v)
        s = O.sf_integral(tau, sfm)
        A = tri_area(v)
        assert np.abs(s - np.array([0, 0, 0, A / 3, A / 3, A / 3])).max() < 1e-13 * max(A, 1.0)


@needs_ref
def test_sf_integral_matches_reference_build():
    rng = np.random.default_rng(2)
    for _ in range(100):
        v = rng.normal(size=(3, 3)) * rng.uniform(0.1, 10.0)
        sfm, rt, tau, cp = REF.el_geom(v)
        a, b = O.sf_integral(tau, sfm), REF.sf_integral(tau, sfm)
        # the vertex entries are cancellation noise (exact value 0) in both
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


# ---------------------------------------------------------------- border

@needs_ref
@pytest.mark.parametrize("tip", [0, 1])
def test_vc_matrix_matches_reference_build(tip):
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof = REF.matrix_vc(nods, conn, 1.0, 0.35, tip_type=tip)
    n = Avc.shape[0] - 1
    # the n x n block is the reference's own make_3dbem_matrix_s (pinned
    # against the oracle's assembly in test_oracle.py)
    A, dof_r = REF.assemble(nods, conn, 1.0, 0.35, tip_type=tip)
    assert (dof == dof_r).all()
    assert (Avc[:n, :n] == A).all()
    row, col = O.vc_border(nods, conn, dof)
    assert np.abs(Avc[n, :n] - row).max() <= 1e-13 * np.abs(row).max()
    assert np.abs(Avc[:n, n] - col).max() == 0.0
    assert Avc[n, n] == 0.0


@pytest.mark.parametrize("stem", ["pennymesh24el_32", "elliptical_440_el_32"])
def test_vc_border_volume_of_uniform_opening(stem):
    """A unit opening along each element's normal (DD = n at all six nodes)
    opens a volume equal to the crack's area."""
    nods, conn = load_mesh(stem)
    n_dof, dof = O.dof_h(nods, conn, 0)
    row, col = O.vc_border(nods, conn, dof)
    d = np.zeros(n_dof)
    for e in range(conn.shape[1]):
        for ns in range(6):
            for j in range(3):
                d[dof[e, 3 * ns + j]] = -col[dof[e, 3 * ns + j]]   # col = -n
    assert abs(row @ d - crack_areas(nods, conn).sum()) < 1e-12 * crack_areas(nods, conn).sum()
    assert np.allclose(np.linalg.norm(col.reshape(-1, 3), axis=1), 1.0, atol=1e-14)


# ---------------------------------------------------------------- truncation

def _vc_matrix(nods, conn, tip):
    A, dof = O.assemble(nods, conn, 1.0, 0.35, tip_type=tip)
    row, col = O.vc_border(nods, conn, dof)
    n = A.shape[0]
    Avc = np.zeros((n + 1, n + 1), order="F")
    Avc[:n, :n], Avc[n, :n], Avc[:n, n] = A, row, col
    return Avc, dof


def test_mod_system_vc_is_selection():
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof0 = _vc_matrix(nods, conn, 0)
    n0 = Avc.shape[0] - 1
    n1, dof1 = O.dof_h(nods, conn, 1)
    rng = np.random.default_rng(3)
    dt = rng.normal(size=n1)
    M, rhs = O.mod_system_vc(Avc, dof0, dof1, n1, dt, 0.25)
    P = incidence(dof1, dof0, n1, n0)
    assert np.abs(M[:n1, :n1] - P.T @ Avc[:n0, :n0] @ P).max() == 0.0
    assert np.abs(M[n1, :n1] - Avc[n0, :n0] @ P).max() == 0.0
    assert np.abs(M[:n1, n1] - P.T @ Avc[:n0, n0]).max() == 0.0
    assert M[n1, n1] == 0.0
    assert (rhs[:n1] == dt).all() and rhs[n1] == 0.25
    # identity truncation gives the matrix back
    M0, rhs0 = O.mod_system_vc(Avc, dof0, dof0, n0, np.zeros(n0), 1.0)
    assert (M0 == Avc).all()


def test_mod_system_vc_sums_merged_dof():
    """Two slots mapped to one used DoF add their rows and columns."""
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof0 = _vc_matrix(nods, conn, 0)
    n0 = Avc.shape[0] - 1
    dof1 = dof0.copy()
    dof1[dof1 == n0 - 1] = 5               # last DoF merged into DoF 5
    n1 = n0 - 1
    full = np.arange(dof0.size, dtype=float)
    M, rhs = O.mod_system_vc(Avc, dof0, dof1, n1, full, -1.0)
    P = incidence(dof1, dof0, n1, n0)
    assert np.allclose(M[:n1, :n1], P.T @ Avc[:n0, :n0] @ P, rtol=0, atol=1e-15 * np.abs(Avc).max())
    assert np.allclose(M[n1, :n1], Avc[n0, :n0] @ P, rtol=0, atol=1e-18)
    # rhs with a full-length delta_t: the last slot in element order wins
    e, j = np.argwhere(dof1 == 5)[-1]
    assert rhs[5] == full[e * 18 + j]


def test_mod_system_vc_rejects_invalid():
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof0 = _vc_matrix(nods, conn, 1)
    n0 = Avc.shape[0] - 1
    _, dof_all = O.dof_h(nods, conn, 0)
    with pytest.raises(ValueError):            # used slot fixed in the original
        O.mod_system_vc(Avc, dof0, dof_all, dof_all.max() + 1, np.zeros(dof_all.max() + 1), 0.0)
    with pytest.raises(ValueError):            # delta_t of no valid length
        O.mod_system_vc(Avc, dof0, dof0, n0, np.zeros(n0 + 3), 0.0)


@needs_ref
@pytest.mark.parametrize("tsize", ["full", "used"])
def test_mod_system_vc_matches_reference_build(tsize):
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof0 = REF.matrix_vc(nods, conn, 1.0, 0.35, tip_type=0)
    n1, dof1 = O.dof_h(nods, conn, 1)
    rng = np.random.default_rng(4)
    dt = rng.normal(size=dof0.size if tsize == "full" else n1)
    Mr, rr = REF.mod_system_vc(Avc, dof0, dof1, n1, dt, 0.5)
    Mo, ro = O.mod_system_vc(Avc, dof0, dof1, n1, dt, 0.5)
    assert (Mr == Mo).all() and (rr == ro).all()


def test_vc_solution_is_scaled_pressure_solution():
    """Prescribed volume, zero traction change: A d + c p = 0, V.d = dV with
    the border column c = -e3 (e3 the element normal).  The far-field load
    sigma_zz = 1 on the planar penny gives the right-hand side e3 (traction
    -sigma.n at the collocation normal n = -e3), A d1 = e3, so d = p d1 with
    p = dV / (V.d1); V.d1 is Sneddon's volume 8 (1 - nu) a^3 / (3 G) up to
    the discretisation error."""
    nods, conn = load_mesh("pennymesh121el_32")
    Avc, dof = _vc_matrix(nods, conn, 1)
    n = Avc.shape[0] - 1
    r = O.rhs(nods, conn, dof, np.array([0, 0, 1.0, 0, 0, 0]))
    LU, piv, info = O.getrf(Avc[:n, :n])
    d1 = O.getrs(LU, piv, r)
    b = np.zeros(n + 1)
    b[n] = 0.3
    LUv, pivv, info = O.getrf(Avc)
    assert info == 0
    x = O.getrs(LUv, pivv, b)
    V1 = Avc[n, :n] @ d1
    p = 0.3 / V1
    assert abs(x[n] - p) < 1e-10 * abs(p)
    assert np.abs(x[:n] - p * d1).max() < 1e-10 * np.abs(x[:n]).max()
    assert abs(Avc[n, :n] @ x[:n] - 0.3) < 1e-12
    assert abs(V1 / (8 * 0.65 / 3) - 1) < 0.1
