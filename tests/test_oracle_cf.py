"""Pins the oracle's step of the volume-control cohesion / friction iteration
(hfpo_vc_cf_iteration, restating src/Development/c_f_iteration.cpp:27-436)
before it checks the HIP path.  CPU only.

The reference file does not build here (il::linearSolve needs LAPACKE), so
the step is pinned by what the mathematics fixes:
  * full separation (zero cohesion, tensile far field, state at rest): the
    step is the plain volume-control solve [A c; v 0][d; p] = [b; V], also
    with tip DoF fixed and with DD in element frames;
  * the step is a Newton step on its active set: from that solution with
    the crack at the new pressure everywhere, the next step moves nothing
    and its residual vanishes;
  * intact crack (compressive far field, no shear): every node leaves the
    system, nothing moves, the residual is 0;
  * slip (compressive far field with shear, friction, no cohesion): after
    the step the shear traction of every node equals the friction bound
    mu |sigma_n| along its shear direction, recomputed from the new DD with
    the matrix and the far field;
  * rotation: with DD in element frames the slip step on the penny rotated
    out of z = 0 also reaches the friction bound in one step (the next
    step's residual vanishes); with global DD it does not, because the laws
    read the opening and slip from the DD components as given (the
    reference's rotation to element frames is an empty branch,
    cohesion_friction.cpp:50-63);
  * invalid handles are rejected.
"""
import numpy as np
import pytest

import oracle_lib
from test_oracle import load_mesh

O = oracle_lib.oracle()
G, NU, BETA = 1.0, 0.3, 0.125


def vc_system(nods, conn, tip=0, dd_local=False):
    A, dof = O.assemble(nods, conn, G, NU, beta=BETA, tip_type=tip, dd_local=dd_local)
    row, col = O.vc_border(nods, conn, dof, dd_local=dd_local)
    n = A.shape[0]
    Avc = np.zeros((n + 1, n + 1), order="F")
    Avc[:n, :n], Avc[n, :n], Avc[:n, n] = A, row, col
    return Avc, dof


def material(cr_open=1e-9, peak_ts=0.0, cr_slip=1.0, peak_sc=0.0, mu=0.0):
    """One surface material: box cohesion, constant friction mu."""
    par = np.array([[cr_open, peak_ts, cr_slip, peak_sc, mu, mu]])
    cmid = np.zeros((1, 2), dtype=np.int32)
    return par, cmid


def frames(nods, conn):
    out = []
    for e in range(conn.shape[1]):
        _, rt, _, _ = O.el_geom(nods[:3, conn[:3, e]], BETA)
        out.append(rt)
    return out


def step(nods, conn, Avc, dof, par, cmid, s_inf, new_vol, dd=None, pp=None, cur=None, dd_local=False,
         mr_open=None, mr_slip=None):
    n_el = conn.shape[1]
    n_nod = 6 * n_el
    return O.vc_cf_iteration(nods, conn, dof, Avc, np.zeros(n_nod, np.int32), cmid, par,
                             np.zeros(n_nod) if mr_open is None else mr_open,
                             np.zeros(n_nod) if mr_slip is None else mr_slip, s_inf, (0, 0), new_vol,
                             np.zeros((n_nod, 3)) if dd is None else dd, np.zeros(n_nod) if pp is None else pp,
                             dof if cur is None else cur, beta=BETA, dd_local=dd_local)


def plain_vc(nods, conn, Avc, dof, s_inf, new_vol):
    n = Avc.shape[0] - 1
    b = np.zeros(n + 1)
    b[:n] = O.rhs(nods, conn, dof, s_inf)
    b[n] = new_vol
    x = np.linalg.solve(Avc, b)
    return O.write_dd(x[:n], dof), x[n], x


@pytest.mark.parametrize("tip,dd_local", [(0, False), (1, False), (0, True)])
def test_full_separation_is_the_plain_vc_solve(tip, dd_local):
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof = vc_system(nods, conn, tip, dd_local)
    n = Avc.shape[0] - 1
    s_inf = np.array([0.0, 0.0, 1.0, 0.0, 0.0, 0.0])
    par, cmid = material()
    # twice the volume the far field alone opens: positive pressure
    x_s = np.linalg.solve(Avc[:n, :n], O.rhs(nods, conn, dof, s_inf))
    vol = 2.0 * Avc[n, :n] @ x_s
    r = step(nods, conn, Avc, dof, par, cmid, s_inf, vol, dd_local=dd_local)
    dd_ref, p_ref, x_ref = plain_vc(nods, conn, Avc, dof, s_inf, vol)
    assert r["rc"] == 0 and r["n_dof"] == n
    assert np.array_equal(r["dof"], dof)
    assert p_ref > 0
    assert abs(r["pressure"] - p_ref) < 1e-11 * abs(p_ref)
    assert np.abs(r["dd"] - dd_ref).max() < 1e-11 * np.abs(dd_ref).max()
    assert np.abs(r["incr"] - x_ref).max() < 1e-11 * np.abs(x_ref).max()
    # the right-hand side by original DoF is -t_inf
    assert np.abs(r["delta_trac"] - O.rhs(nods, conn, dof, s_inf)).max() < 1e-14
    # every collocation point opened: the pressure is applied everywhere
    assert np.abs(r["pp"] - r["pressure"]).max() == 0.0


def test_step_from_its_own_solution_moves_nothing():
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof = vc_system(nods, conn)
    n = Avc.shape[0] - 1
    s_inf = np.array([0.0, 0.0, 0.8, 0.0, 0.0, 0.0])
    par, cmid = material()
    vol = 1.5 * Avc[n, :n] @ np.linalg.solve(Avc[:n, :n], O.rhs(nods, conn, dof, s_inf))
    r1 = step(nods, conn, Avc, dof, par, cmid, s_inf, vol)
    r2 = step(nods, conn, Avc, dof, par, cmid, s_inf, vol, dd=r1["dd"], pp=r1["pp"], cur=r1["dof"])
    # every node is at zero effective traction: with zero cohesion the laws
    # split the nodes on the sign of rounding noise (nt <= 0 and no shear is
    # "intact"), so the active set is not pinned here, only that nothing moves
    assert r2["rc"] == 0 and 0 < r2["n_dof"] <= n
    assert r2["residual"] < 1e-10 * r1["residual"]
    assert np.abs(r2["dd"] - r1["dd"]).max() < 1e-12 * np.abs(r1["dd"]).max()
    assert abs(r2["pressure"] - r1["pressure"]) < 1e-12 * abs(r1["pressure"])


@pytest.mark.parametrize("mu", [0.0, 0.4])
def test_intact_crack_leaves_the_system(mu):
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof = vc_system(nods, conn)
    par, cmid = material(mu=mu)
    s_inf = np.array([0.0, 0.0, -1.0, 0.0, 0.0, 0.0])
    pp = np.full(6 * conn.shape[1], 0.25)
    r = step(nods, conn, Avc, dof, par, cmid, s_inf, 0.1, pp=pp)
    assert r["rc"] == 1 and r["n_dof"] == 0
    assert np.all(r["dof"] == -1)
    assert np.all(r["dd"] == 0.0)
    # closed nodes keep their pressure; nothing is adjusted
    assert np.array_equal(r["pp"], pp) and r["pressure"] == 0.25
    assert r["residual"] == 0.0


def test_slip_reaches_the_friction_bound():
    nods, conn = load_mesh("pennymesh121el_32")
    Avc, dof = vc_system(nods, conn)
    n = Avc.shape[0] - 1
    n_el = conn.shape[1]
    mu = 0.2
    par, cmid = material(mu=mu)
    # sigma_zz = -1 (compression), sigma_xz = 0.3, sigma_yz = -0.1: shear
    # traction 0.316 > mu |sigma_n| = 0.2 on the whole crack plane
    s_inf = np.array([0.0, 0.0, -1.0, 0.0, 0.3, -0.1])
    r = step(nods, conn, Avc, dof, par, cmid, s_inf, 0.0)
    assert r["rc"] == 0 and r["n_dof"] == n  # shear slips, the normal DoF stay as they were
    # tractions of the new DD and the far field, per collocation node, in the
    # element frame: the shear imbalance st = -(local shear traction)
    b = O.rhs(nods, conn, dof, s_inf)  # = -t_inf
    x = np.zeros(n)
    for e in range(n_el):
        for q in range(6):
            for i in range(3):
                x[dof[e, 3 * q + i]] = r["dd"][6 * e + q, i]
    t_tot = Avc[:n, :n] @ x - b
    R = frames(nods, conn)
    worst = 0.0
    for e in range(n_el):
        for q in range(6):
            t = R[e] @ t_tot[dof[e, 3 * q:3 * q + 3]]
            st = -t[:2]
            ti = R[e] @ (-b[dof[e, 3 * q:3 * q + 3]])
            st0 = -ti[:2]  # the imbalance at rest: the far field's shear
            nt0 = -ti[2]
            bound = -mu * nt0
            target = bound * st0 / np.linalg.norm(st0)
            worst = max(worst, np.abs(st - target).max())
    assert worst < 1e-10


def test_rejects_handle_outside_the_original():
    nods, conn = load_mesh("pennymesh24el_32")
    Avc, dof = vc_system(nods, conn, tip=1)
    par, cmid = material()
    cur = dof.copy()
    fixed = np.argwhere(dof < 0)[0]
    cur[tuple(fixed)] = 0
    with pytest.raises(ValueError):
        step(nods, conn, Avc, dof, par, cmid, np.array([0, 0, 1.0, 0, 0, 0]), 0.1, cur=cur)


def rotated(nods, R):
    out = nods.copy()
    out[:3] = R @ nods[:3]
    return out


@pytest.mark.parametrize("dd_local", [True, False])
def test_slip_on_a_rotated_crack(dd_local):
    nods, conn = load_mesh("pennymesh121el_32")
    c, s = np.cos(np.pi / 6), np.sin(np.pi / 6)
    R = np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])
    nods = rotated(nods, R)
    Avc, dof = vc_system(nods, conn, dd_local=dd_local)
    par, cmid = material(mu=0.2)
    nrm, t = R @ np.array([0, 0, 1.0]), R @ np.array([1.0, 0, 0])
    S = -np.outer(nrm, nrm) + 0.3 * (np.outer(nrm, t) + np.outer(t, nrm))
    s_inf = np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])
    n_nod = 6 * conn.shape[1]
    ms = np.full(n_nod, 0.5)
    r1 = step(nods, conn, Avc, dof, par, cmid, s_inf, 0.0, dd_local=dd_local, mr_slip=ms)
    r2 = step(nods, conn, Avc, dof, par, cmid, s_inf, 0.0, dd=r1["dd"], pp=r1["pp"], cur=r1["dof"],
              dd_local=dd_local, mr_slip=ms)
    if dd_local:
        assert r2["residual"] < 1e-12 * r1["residual"]
    else:
        assert r2["residual"] > 0.1 * r1["residual"]
