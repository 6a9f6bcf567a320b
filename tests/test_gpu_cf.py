"""One step of the volume-control cohesion / friction iteration
(hfpx_vc_cf_iteration, through the C ABI via the vc_cf_iteration mirror)
against the oracle's step (GPU).  The oracle's step is pinned in
test_oracle_cf.py (the plain volume-control solve, the Newton fixed point,
the friction bound).

Each comparison starts both from the same state.  The states are chosen so
that no node sits on a decision threshold of the laws to rounding (such a
node would be classified by rounding noise in either implementation)."""
import numpy as np
import pytest

import oracle_lib
from oracle_lib import golden

pytestmark = pytest.mark.gpu

O = oracle_lib.oracle()
G, NU, BETA = 1.0, 0.3, 0.125


def load_mesh(stem):
    nods = np.load(golden(f"Nodes_{stem}.npy"))
    conn = np.load(golden(f"Elems_{stem}.npy")).astype(np.int64)
    conn[:3] -= 1
    return nods, conn


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def hx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    from hfpx3d_amd import hfp3d
    return hfp3d


class Case:
    """A mesh with its volume-control matrix on the device (the mirror's
    SAE_T) and in the oracle, and a surface material."""

    def __init__(self, hx, stem, tip=0, dd_local=False, par=None):
        self.hx = hx
        self.nods, self.conn = load_mesh(stem)
        self.n_el = self.conn.shape[1]
        self.n_nod = 6 * self.n_el
        self.dd_local = dd_local
        mesh = hx.Mesh_Geom_T(nods=self.nods, conn=self.conn)
        self.n_par = hx.Num_Param_T(beta=BETA, tip_type=tip, is_dd_local=dd_local)
        self.dof = hx.DoF_Handle_T()
        A64 = hx.make_3dbem_matrix_vc(G, NU, mesh, self.n_par, self.dof)
        self.sys = hx.SAE_T(n_dof=self.dof.n_dof, matrix=A64)
        self.mesh = mesh
        A, dof_o = O.assemble(self.nods, self.conn, G, NU, beta=BETA, tip_type=tip, dd_local=dd_local)
        assert np.array_equal(dof_o, self.dof.dof_h)
        row, col = O.vc_border(self.nods, self.conn, dof_o, dd_local=dd_local)
        n = A.shape[0]
        self.Avc = np.zeros((n + 1, n + 1), order="F")
        self.Avc[:n, :n], self.Avc[n, :n], self.Avc[:n, n] = A, row, col
        self.n0 = n
        # one material: box cohesion (0), constant friction (0)
        self.par = np.array([[1e-9, 0.0, 1.0, 0.0, 0.0, 0.0]]) if par is None else np.atleast_2d(par)
        self.cmid = np.zeros((self.par.shape[0], 2), dtype=np.int32)

    def prop(self):
        hx = self.hx
        return hx.Properties_T(c_model_id=self.cmid,
                               f_c_param=[hx.F_C_Param_T(*row) for row in self.par])

    def run_both(self, s_inf, new_vol, dd, pp, cur, mr_open, mr_slip, mat_id=None, inj=(0, 0)):
        hx = self.hx
        mat_id = np.zeros(self.n_nod, np.int32) if mat_id is None else mat_id
        o = O.vc_cf_iteration(self.nods, self.conn, self.dof.dof_h, self.Avc, mat_id, self.cmid, self.par, mr_open,
                              mr_slip, s_inf, inj, new_vol, dd, pp, cur, beta=BETA, dd_local=self.dd_local)
        md = hx.Mesh_Data_T(mesh=self.mesh, dof_h_dd=hx.DoF_Handle_T(dof_h=cur.copy(), n_dof=int(cur.max()) + 1),
                            dd=dd.copy(), pp=pp.copy(), mat_id=mat_id,
                            frac_state=hx.Frac_State_T(mr_open.copy(), mr_slip.copy(), None, None, None))
        rep = hx._lib.CfReport()
        res = hx.vc_cf_iteration(None, self.sys, self.dof, self.prop(), self.n_par,
                                 hx.Load_T(s_inf=np.asarray(s_inf, float), inj_loc=np.array([inj])), new_vol, md,
                                 tol=1e-14, report=rep)
        return o, md, res, rep

    def compare(self, o, md, res, rep, tol=1e-9, res_atol=0.0):
        assert np.array_equal(md.dof_h_dd.dof_h, o["dof"])
        assert md.dof_h_dd.n_dof == o["n_dof"] == rep.n_dof
        assert rel(md.dd, o["dd"]) < tol
        assert rel(md.pp, o["pp"]) < tol if np.abs(o["pp"]).max() > 0 else np.all(md.pp == 0)
        assert abs(rep.pressure - o["pressure"]) <= tol * max(abs(o["pressure"]), 1e-300)
        assert abs(res - o["residual"]) <= tol * o["residual"] + res_atol
        fs = md.frac_state
        for k in ("friction_coef", "slip_cohesion", "open_cohesion"):
            assert np.abs(getattr(fs, k) - o[k]).max() <= 1e-14 * max(1.0, np.abs(o[k]).max())


@pytest.mark.parametrize("stem,tip,dd_local", [("pennymesh121el_32", 0, False), ("pennymesh121el_32", 1, False),
                                               ("pennymesh24el_32", 0, True)])
def test_full_separation_matches_oracle(hx, stem, tip, dd_local):
    print(f"case: cf full separation {stem} tip {tip} local {dd_local}", flush=True)
    c = Case(hx, stem, tip, dd_local)
    s_inf = np.array([0.0, 0.0, 1.0, 0.0, 0.0, 0.0])
    z = np.zeros(c.n_nod)
    o, md, res, rep = c.run_both(s_inf, 0.5, np.zeros((c.n_nod, 3)), z, c.dof.dof_h.copy(), z, z)
    assert o["rc"] == 0 and rep.refactored == 1 and rep.ir.converged == 1
    c.compare(o, md, res, rep)


def test_slip_steps_reuse_the_factors(hx):
    """Compression with shear above the friction bound, slip already started
    (mr_slip > 0): every shear DoF slips, every normal DoF stays, so the
    active set is the same at every step and the FP32 factors of the first
    step serve the next ones.  Each step is checked against the oracle from
    the same state."""
    print("case: cf slip steps", flush=True)
    c = Case(hx, "pennymesh121el_32", par=[1e-9, 0.0, 1.0, 0.0, 0.2, 0.2])
    s_inf = np.array([0.0, 0.0, -1.0, 0.0, 0.3, -0.1])
    z = np.zeros(c.n_nod)
    dd, pp, cur = np.zeros((c.n_nod, 3)), z.copy(), c.dof.dof_h.copy()
    mr_slip = np.full(c.n_nod, 0.5)
    res0 = None
    for k in range(3):
        o, md, res, rep = c.run_both(s_inf, 0.0, dd, pp, cur, z, mr_slip)
        assert o["n_dof"] == c.n0
        assert rep.refactored == (1 if k == 0 else 0)
        # from the second step on the state is the fixed point: the residual
        # is rounding noise, compared against the first step's scale
        c.compare(o, md, res, rep, res_atol=0.0 if res0 is None else 1e-12 * res0)
        if res0 is None:
            res0 = o["residual"]
        else:
            assert res < 1e-12 * res0
        dd, pp, cur = md.dd, md.pp, md.dof_h_dd.dof_h


def node_xy(nods, conn):
    """In-plane position of every node (6 per element: vertices, then the
    mid-edge node opposite each vertex)."""
    n_el = conn.shape[1]
    xy = np.zeros((6 * n_el, 2))
    for e in range(n_el):
        v = nods[:2, conn[:3, e]]
        for q in range(6):
            xy[6 * e + q] = v[:, q] if q < 3 else 0.5 * (v[:, (q - 2) % 3] + v[:, (q - 1) % 3])
    return xy


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mixed_states_match_oracle(hx, seed):
    """Compression with a little shear on the planar penny; the half x > 0 is
    open with fluid at pressure 0.5 (a smooth opening below the critical
    one, perturbed by 5 % noise), fully broken beyond x = 0.5, slipping
    beyond y = 0.5; the rest is closed.  The laws put the nodes in every
    branch: cohesive zone, full separation, slip, intact.  (Perturbing the
    state by 1e-11 leaves the oracle's new handle unchanged for these
    seeds.)"""
    print(f"case: cf mixed state seed {seed}", flush=True)
    rng = np.random.default_rng(seed)
    c = Case(hx, "pennymesh121el_32", par=[0.02, 0.3, 0.05, 0.1, 0.4, 0.4])
    n = c.n_nod
    xy = node_xy(c.nods, c.conn)
    r = np.hypot(xy[:, 0], xy[:, 1])
    dd = np.zeros((n, 3))
    op = 0.01 * np.sqrt(np.clip(1 - r ** 2, 0, None)) * (1 + 0.3 * xy[:, 0])
    dd[:, 2] = np.where(xy[:, 0] > 0, op, 0.0) * (1 + 0.05 * rng.standard_normal(n))
    pp = np.where(xy[:, 0] > 0, 0.5, 0.0) + 0.01 * rng.random(n)
    mr_open = np.where(xy[:, 0] > 0.5, 1.0, 0.0)
    mr_slip = np.where(xy[:, 1] > 0.5, 0.5, 0.0)
    s_inf = np.array([0.0, 0.0, -0.3, 0.0, 0.05, 0.0])
    o, md, res, rep = c.run_both(s_inf, 0.02, dd, pp, c.dof.dof_h.copy(), mr_open, mr_slip, inj=(3, 4))
    assert 0 < o["n_dof"] < c.n0  # some nodes left the system, some stayed
    c.compare(o, md, res, rep)


def test_intact_crack_moves_nothing(hx):
    print("case: cf intact", flush=True)
    c = Case(hx, "pennymesh24el_32", par=[1e-9, 0.0, 1.0, 0.0, 0.4, 0.4])
    s_inf = np.array([0.0, 0.0, -1.0, 0.0, 0.0, 0.0])
    z = np.zeros(c.n_nod)
    pp = np.full(c.n_nod, 0.25)
    o, md, res, rep = c.run_both(s_inf, 0.1, np.zeros((c.n_nod, 3)), pp, c.dof.dof_h.copy(), z, z)
    assert o["n_dof"] == rep.n_dof == 0 and rep.refactored == 0
    assert np.all(md.dof_h_dd.dof_h == -1)
    assert np.all(md.dd == 0.0) and np.array_equal(md.pp, pp)
    assert rep.pressure == 0.25 and res == 0.0


def test_rejects_handle_outside_the_original(hx):
    c = Case(hx, "pennymesh24el_32", tip=1)
    z = np.zeros(c.n_nod)
    cur = c.dof.dof_h.copy()
    fixed = np.argwhere(cur < 0)[0]
    cur[tuple(fixed)] = 0
    md = hx.Mesh_Data_T(mesh=c.mesh, dof_h_dd=hx.DoF_Handle_T(dof_h=cur, n_dof=c.n0), dd=np.zeros((c.n_nod, 3)),
                        pp=z.copy(), mat_id=np.zeros(c.n_nod, np.int32), frac_state=hx.Frac_State_T.zeros(c.n_nod))
    with pytest.raises(hx._lib.HfpxError):
        hx.vc_cf_iteration(None, c.sys, c.dof, c.prop(), c.n_par,
                           hx.Load_T(s_inf=np.array([0, 0, 1.0, 0, 0, 0]), inj_loc=np.array([[0, 0]])), 0.1, md)


def test_cpp_host_mirror_matches_oracle():
    """Two steps of hfp3d::vc_cf_iteration of the C++ host mirror (through
    the hfpx3d_cf_check program) against the oracle's steps: full separation
    of the 24-element penny under sigma_zz = 1 with twice the volume the far
    field opens, then the step from that solution."""
    import json
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "native", "hfpx3d_cf_check")
    nods, conn = load_mesh("pennymesh24el_32")
    A, dof = O.assemble(nods, conn, G, NU, beta=BETA, tip_type=0)
    row, col = O.vc_border(nods, conn, dof)
    n = A.shape[0]
    Avc = np.zeros((n + 1, n + 1), order="F")
    Avc[:n, :n], Avc[n, :n], Avc[:n, n] = A, row, col
    vol = 2.0 * row @ np.linalg.solve(A, O.rhs(nods, conn, dof, np.array([0, 0, 1.0, 0, 0, 0])))
    print("case: C++ host mirror of vc_cf_iteration, 24-element penny", flush=True)
    r = subprocess.run([exe, golden("Elems_pennymesh24el_32.npy"), golden("Nodes_pennymesh24el_32.npy"), "1.0",
                        repr(float(vol))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    n_nod = 6 * conn.shape[1]
    par = np.array([[1e-9, 0.0, 1.0, 0.0, 0.0, 0.0]])
    cmid = np.zeros((1, 2), np.int32)
    z = np.zeros(n_nod)
    dd, pp, cur = np.zeros((n_nod, 3)), z.copy(), dof.copy()
    for k in range(2):
        o = O.vc_cf_iteration(nods, conn, dof, Avc, np.zeros(n_nod, np.int32), cmid, par, z, z,
                              np.array([0, 0, 1.0, 0, 0, 0]), (0, 0), vol, dd, pp, cur, beta=BETA)
        got_dd = np.array(d[f"dd{k}"]).reshape(n_nod, 3, order="F")
        if k == 0:
            assert d["n_dof0"] == o["n_dof"] == n
            assert rel(got_dd, o["dd"]) < 1e-9
            assert rel(np.array(d["pp0"]), o["pp"]) < 1e-9
            assert abs(d["residual0"] - o["residual"]) < 1e-9 * o["residual"]
            scale = o["residual"]
        else:
            # from its own solution nothing moves (the active set there is
            # decided by rounding noise, so it is not compared)
            assert d["residual1"] < 1e-9 * scale and o["residual"] < 1e-9 * scale
            assert rel(got_dd, o["dd"]) < 1e-9
        dd, pp, cur = o["dd"], o["pp"], o["dof"]
