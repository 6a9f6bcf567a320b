"""Non-uniform edge partition (per-node vertex weights) in the oracle (CPU).

The oracle's shape functions and collocation points for vertex weights are
pinned
  * against the reference build's make_el_sfm_nonuniform / el_cp_nonuniform
    (src/Core/element_utilities.cpp:240-322, 397-431), where it is built;
  * by the Lagrange property N_j(node_k) = delta_jk at the weighted nodes,
    the partition of unity, and equal weights giving the uniform element.
The oracle's kernel with a non-uniform basis is pinned against the
reference's make_local_3dbem_submatrix fed the reference's own basis, and
the weighted assembly, stresses and volume-control border by linearity:
both bases span the same quadratic polynomials, so a DD field on the
weighted basis equals the uniform-basis field with the values at the
uniform nodes, and its stresses and tractions must agree."""
import numpy as np
import pytest

import oracle_lib
from oracle_lib import golden

O = oracle_lib.oracle()
REF = oracle_lib.reference_or_none()


def load_mesh(stem):
    nods = np.load(golden(f"Nodes_{stem}.npy"))
    conn = np.load(golden(f"Elems_{stem}.npy")).astype(np.int64)
    conn[:3] -= 1
    return nods, conn


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def eval_sfm(sfm, t):
    """Shape-function values (6) at the complex local coordinate t."""
    mono = np.array([1, t, np.conj(t), t * t, np.conj(t) ** 2, t * np.conj(t)])
    return (sfm @ mono).real


def random_elements(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        v = rng.normal(size=(3, 3))
        w = rng.uniform(0.25, 4.0, 3)
        out.append((v, w))
    return out


def local_nodes(tau, w):
    """Complex local coordinates of the six nodes for vertex weights w."""
    nodes = list(tau)
    for v in range(3):
        a, b = (v + 1) % 3, (v + 2) % 3
        nodes.append((w[a] * tau[a] + w[b] * tau[b]) / (w[a] + w[b]))
    return nodes


@pytest.mark.skipif(REF is None, reason="reference build (oracle/_ref) not available")
def test_nonuniform_geometry_matches_reference_build():
    for v, w in random_elements(20, 1):
        for beta in (0.0, 0.125, 0.3):
            sfm, rt, tau, cp = O.el_geom(v, beta, wts=w)
            sr, rr, cr = REF.el_geom_nonuniform(v, w, beta)
            assert rel(sfm, sr) < 1e-13
            assert np.abs(rt - rr).max() < 1e-14
            assert np.abs(cp - cr).max() < 1e-14


def test_nonuniform_basis_is_lagrange():
    for v, w in random_elements(20, 2):
        sfm, rt, tau, cp = O.el_geom(v, 0.0, wts=w)
        vals = np.array([eval_sfm(sfm, t) for t in local_nodes(tau, w)])  # row k = node k
        assert np.abs(vals - np.eye(6)).max() < 1e-12
        # partition of unity: the rows sum to the constant 1
        s = sfm.sum(axis=0)
        assert abs(s[0] - 1) < 1e-12 and np.abs(s[1:]).max() < 1e-11
        # collocation points at beta = 0 are the nodes (global frame)
        for k, t in enumerate(local_nodes(tau, w)):
            x = v[:, 0] + rt.T @ np.array([t.real, t.imag, 0.0])
            assert np.abs(cp[:, k] - x).max() < 1e-12


def test_equal_weights_give_the_uniform_element():
    for v, w in random_elements(10, 3):
        for c in (1.0, 2.5):
            sfm_w, _, _, cp_w = O.el_geom(v, 0.125, wts=[c, c, c])
            sfm_u, _, _, cp_u = O.el_geom(v, 0.125)
            assert rel(sfm_w, sfm_u) < 1e-14
            assert np.abs(cp_w - cp_u).max() < 1e-14


@pytest.mark.skipif(REF is None, reason="reference build (oracle/_ref) not available")
def test_kernel_with_nonuniform_basis_matches_reference_build():
    rng = np.random.default_rng(4)
    for v, w in random_elements(6, 5):
        sr, rr, _ = REF.el_geom_nonuniform(v, w, 0.125)
        _, _, tau, _ = REF.el_geom(v, 0.125)
        for _ in range(4):
            h = rng.choice([0.0, rng.normal(scale=0.5)])
            z = complex(*rng.normal(size=2))
            Sr = REF.local_submatrix(1.0, 0.3, h, z, tau, sr)
            So = O.local_submatrix(1.0, 0.3, h, z, tau, sr)
            assert rel(So, Sr) < 1e-10


def basis_change(sfm_w, tau):
    """T (6x6): row j = values of the weighted N_j at the uniform nodes, so a
    weighted-basis DD d_w gives the uniform-basis DD T^T d_w."""
    uni = local_nodes(tau, [1.0, 1.0, 1.0])
    return np.array([[eval_sfm(sfm_w, t)[j] for t in uni] for j in range(6)])


def weighted_mesh(seed):
    nods, conn = load_mesh("pennymesh24el_32")
    rng = np.random.default_rng(seed)
    # lift the flat mesh so that off-plane branches are exercised too
    nods = nods.copy()
    nods[2] = 0.05 * np.sin(2.0 * nods[0]) * np.cos(1.5 * nods[1])
    wts = rng.uniform(0.4, 2.5, nods.shape[1])
    return nods, conn, wts


def test_weighted_stress_equals_uniform_stress_of_the_same_field():
    nods, conn, wts = weighted_mesh(6)
    n_el = conn.shape[1]
    rng = np.random.default_rng(7)
    dd_w = rng.normal(size=(6 * n_el, 3))
    dd_u = np.zeros_like(dd_w)
    for e in range(n_el):
        v = nods[:3, conn[:3, e]]
        sfm, _, tau, _ = O.el_geom(v, 0.0, wts=wts[conn[:3, e]])
        T = basis_change(sfm, tau)
        dd_u[6 * e:6 * e + 6] = T.T @ dd_w[6 * e:6 * e + 6]
    pts = rng.uniform(-1.5, 1.5, size=(3, 40))
    s_inf = np.zeros(6)
    Sw = O.stress(nods, conn, 1.0, 0.3, dd_w, pts, s_inf, node_wts=wts)
    Su = O.stress(nods, conn, 1.0, 0.3, dd_u, pts, s_inf)
    assert rel(Sw, Su) < 1e-11


def test_weighted_assembly_equals_uniform_tractions_at_weighted_points():
    """Column (source e, node j, comp c) of the weighted matrix = traction at
    the weighted collocation points of the weighted N_j, i.e. of sum_k
    T_jk N_k on the uniform basis (tractions from the oracle's stresses)."""
    nods, conn, wts = weighted_mesh(8)
    n_el = conn.shape[1]
    beta = 0.125
    A, dof = O.assemble(nods, conn, 1.0, 0.3, beta=beta, node_wts=wts)
    geo = [O.el_geom(nods[:3, conn[:3, e]], beta, wts=wts[conn[:3, e]]) for e in range(n_el)]
    pts = np.concatenate([g[3] for g in geo], axis=1)                          # 3 x 6 n_el
    nrm = np.concatenate([np.repeat(-g[1][2:3].T, 6, axis=1) for g in geo], axis=1)
    z6 = np.zeros(6)
    for es in (0, 5, 17):
        sfm, _, tau, _ = geo[es][0], None, geo[es][2], None
        T = basis_change(geo[es][0], geo[es][2])
        for j in (0, 4):
            for c in (0, 2):
                dd_w = np.zeros((6 * n_el, 3))
                dd_w[6 * es + j, c] = 1.0
                dd_u = np.zeros_like(dd_w)
                dd_u[6 * es:6 * es + 6] = T.T @ dd_w[6 * es:6 * es + 6]
                S = O.stress(nods, conn, 1.0, 0.3, dd_u, pts, z6)               # n_pts x 6
                sig = np.stack([np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]]) for s in S])
                tr = np.einsum("pij,jp->pi", sig, nrm).reshape(-1)             # row 3 cp + i
                col = A[:, dof[es, 3 * j + c]]
                assert rel(col, tr) < 1e-10, (es, j, c)


def test_weighted_volume_row_is_the_area_integral():
    nods, conn, wts = weighted_mesh(9)
    n_el = conn.shape[1]
    n_dof, dof = O.dof_h(nods, conn, 0)
    row, col = O.vc_border(nods, conn, dof, node_wts=wts)
    row_u, col_u = O.vc_border(nods, conn, dof)
    assert np.array_equal(col, col_u)
    # sum over an element's nodes of int N_j dA is its area (partition of unity);
    # the weighted vertex integrals are not zero
    for e in range(n_el):
        v = nods[:3, conn[:3, e]]
        area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]))
        _, rt, _, _ = O.el_geom(v, 0.0)
        n2 = rt[2, 2]
        tot = sum(row[dof[e, 3 * j + 2]] for j in range(6)) / n2
        assert abs(tot - area) < 1e-12 * max(1.0, area)
