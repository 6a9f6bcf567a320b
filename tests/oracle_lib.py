"""ctypes access to the oracle (oracle/liboracle.so, the C restatement) and,
when it has been built, to the reference build (oracle/_ref/libhfpref.so).
Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libhfpref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

_dp = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_int = ctypes.c_int


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _f(a, dt=np.float64):
    return np.asfortranarray(a, dtype=dt)


def _pw(w):
    """Per-node (or per-vertex) weights as a pointer, NULL for uniform."""
    return None if w is None else _p(w)


def _wts(w):
    return None if w is None else np.ascontiguousarray(w, dtype=np.float64)


class _Lib:
    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix

    def fn(self, name, res, args):
        f = getattr(self.lib, self.prefix + name)
        f.restype = res
        f.argtypes = args
        return f


def _geom_unpack(sfm, rt, tau, cp):
    return ((sfm[0::2] + 1j * sfm[1::2]).reshape(6, 6, order="F"), rt.reshape(3, 3, order="F"),
            tau[0::2] + 1j * tau[1::2], cp.reshape(3, 6, order="F"))


def _sub_pack(tau, sfm):
    t = np.zeros(6)
    t[0::2], t[1::2] = tau.real, tau.imag
    s = np.zeros(72)
    sf = np.asarray(sfm).reshape(-1, order="F")
    s[0::2], s[1::2] = sf.real, sf.imag
    return t, s


def _fc_args(dd, mat_id, cmid, par):
    dd = _f(dd)
    mid = np.ascontiguousarray(mat_id, dtype=np.int32)
    cm = _f(cmid, np.int32) if cmid is not None else None
    pr = np.ascontiguousarray(par, dtype=np.float64).reshape(-1, 6)
    return dd, mid, cm, pr


class _FrictionCohesion:
    """match_f_c / update_mrd through a library with the hfpo_ / ref_ layout
    (dd n_nod x 3, mat_id n_nod, cmid n_mat x 2, par n_mat x 6)."""

    def match_f_c(self, dd, mat_id, cmid, par, mr_open, mr_slip):
        dd, mid, cm, pr = _fc_args(dd, mat_id, cmid, par)
        n = dd.shape[0]
        mo = np.ascontiguousarray(mr_open, dtype=np.float64)
        ms = np.ascontiguousarray(mr_slip, dtype=np.float64)
        fc, sc, oc = np.zeros(n), np.zeros(n), np.zeros(n)
        self._mfc(n, _p(dd), n, _p(mid), _p(cm), pr.shape[0], _p(pr), _p(mo), _p(ms), _p(fc), _p(sc), _p(oc))
        return fc, sc, oc

    def update_mrd(self, dd, mat_id, par, mr_open, mr_slip):
        dd, mid, _, pr = _fc_args(dd, mat_id, None, par)
        n = dd.shape[0]
        mo = np.array(mr_open, dtype=np.float64)
        ms = np.array(mr_slip, dtype=np.float64)
        self._mrd(n, _p(dd), n, _p(mid), pr.shape[0], _p(pr), _p(mo), _p(ms))
        return mo, ms


class Oracle(_FrictionCohesion):
    """The C restatement (hfpo_* functions)."""

    def __init__(self, path=ORACLE):
        L = _Lib(path, "hfpo_")
        self._geom = L.fn("el_geom", None, [_dp, _d, _dp, _dp, _dp, _dp])
        self._geomw = L.fn("el_geom_w", None, [_dp, _dp, _d, _dp, _dp, _dp, _dp])
        self._sub = L.fn("local_submatrix", None, [_d, _d, _d, _d, _d, _dp, _dp, _dp])
        self._dof = L.fn("dof_h", _i64, [_dp, _i64, _i64, _dp, _i64, _i64, _int, _dp])
        self._asm = L.fn("assemble_w", _i64, [_dp, _i64, _i64, _dp, _i64, _i64, _dp, _i64, _d, _d, _d,
                                              _dp, _i64, _i64, _int, _int, _dp])
        self._rhs = L.fn("rhs", None, [_dp, _i64, _i64, _dp, _i64, _i64, _dp, _dp, _dp])
        self._rows = L.fn("assemble_rows", None, [_dp, _i64, _i64, _dp, _i64, _i64, _dp, _i64, _d, _d, _d, _i64,
                                                   _i64, _int, _dp])
        self._wdd = L.fn("write_dd", None, [_dp, _dp, _i64, _dp])
        self._stress = L.fn("stress_w", None, [_dp, _i64, _i64, _dp, _i64, _i64, _d, _d, _dp, _dp, _i64,
                                               _dp, _dp, _int, _int, _dp])
        self._getrf = L.fn("getrf", _i64, [_i64, _dp, _i64, _dp, _int])
        self._getrs = L.fn("getrs", None, [_i64, _i64, _dp, _i64, _dp, _dp, _i64])
        self._sned = L.fn("sneddon_dd", None, [_d, _d, _d, _d, _dp, _dp])
        self._sfi = L.fn("sf_integral", None, [_dp, _dp, _dp])
        self._vcb = L.fn("vc_border_w", None, [_dp, _i64, _i64, _dp, _i64, _i64, _dp, _dp, _int, _dp, _dp])
        self._mvc = L.fn("mod_system_vc", _int, [_dp, _i64, _i64, _dp, _dp, _i64, _i64, _i64, _dp, _i64, _d,
                                                  _dp, _dp])
        self._mfc = L.fn("match_f_c", None, [_i64, _dp, _i64, _dp, _dp, _i64, _dp, _dp, _dp, _dp, _dp, _dp])
        self._mrd = L.fn("update_mrd", None, [_i64, _dp, _i64, _dp, _i64, _dp, _dp, _dp])
        self._cfi = L.fn("vc_cf_iteration", _int,
                         [_dp, _i64, _i64, _dp, _i64, _i64, _d, _int, _dp, _i64, _dp, _i64, _dp, _dp, _i64, _dp,
                          _dp, _dp, _dp, _i64, _i64, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp])

    def el_geom(self, vert, beta=0.125, wts=None):
        """wts: the three vertex weights of a non-uniform edge partition."""
        v = _f(vert)
        rt, tau, sfm, cp = np.zeros(9), np.zeros(6), np.zeros(72), np.zeros(18)
        if wts is None:
            self._geom(_p(v), beta, _p(rt), _p(tau), _p(sfm), _p(cp))
        else:
            w = _wts(wts)
            self._geomw(_p(v), _p(w), beta, _p(rt), _p(tau), _p(sfm), _p(cp))
        return _geom_unpack(sfm, rt, tau, cp)

    def local_submatrix(self, G, nu, h, z, tau, sfm):
        t, s = _sub_pack(tau, sfm)
        out = np.zeros(108)
        self._sub(G, nu, h, z.real, z.imag, _p(t), _p(s), _p(out))
        return out.reshape(6, 18, order="F")

    def dof_h(self, nods, conn, tip_type):
        nods, conn = _f(nods), _f(conn, np.int64)
        n_el = conn.shape[1]
        dof = np.zeros(18 * n_el, dtype=np.int64)
        n = self._dof(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], n_el, tip_type, _p(dof))
        return int(n), dof.reshape(n_el, 18, order="F")

    def assemble(self, nods, conn, G, nu, beta=0.125, tip_type=0, targets=None, nthreads=8, dd_local=False,
                 node_wts=None):
        nods, conn = _f(nods), _f(conn, np.int64)
        w = _wts(node_wts)
        n_el = conn.shape[1]
        n_dof, dof = self.dof_h(nods, conn, tip_type)
        dof_f = _f(dof, np.int64)
        A = np.zeros(n_dof * n_dof)
        t0, t1 = (0, 0) if targets is None else targets
        self._asm(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], n_el, _p(dof_f), n_dof,
                  G, nu, beta, _pw(w), t0, t1, int(dd_local), nthreads, _p(A))
        return A.reshape(n_dof, n_dof, order="F"), dof

    def assemble_rows(self, nods, conn, G, nu, t0, t1, beta=0.125, tip_type=0, nthreads=8):
        """Rows 18 t0 .. 18 t1 - 1 of the matrix (local traction slots of the
        target elements t0 .. t1 - 1, all columns) without forming it."""
        nods, conn = _f(nods), _f(conn, np.int64)
        n_el = conn.shape[1]
        n_dof, dof = self.dof_h(nods, conn, tip_type)
        dof_f = _f(dof, np.int64)
        nr = 18 * (t1 - t0)
        R = np.zeros(nr * n_dof)
        self._rows(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], n_el, _p(dof_f), n_dof, G, nu,
                   beta, t0, t1, nthreads, _p(R))
        return R.reshape(nr, n_dof, order="F")

    def rhs(self, nods, conn, dof, s_inf):
        nods, conn = _f(nods), _f(conn, np.int64)
        dof_f = _f(dof, np.int64)
        n_dof = int((dof >= 0).sum())
        r = np.zeros(n_dof)
        s = np.ascontiguousarray(s_inf, dtype=np.float64)
        self._rhs(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], conn.shape[1],
                  _p(dof_f), _p(s), _p(r))
        return r

    def write_dd(self, x, dof):
        dof_f = _f(dof, np.int64)
        n_el = dof.shape[0]
        dd = np.zeros(18 * n_el)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        self._wdd(_p(xx), _p(dof_f), n_el, _p(dd))
        return dd.reshape(6 * n_el, 3, order="F")

    def stress(self, nods, conn, G, nu, dd, pts, s_inf, nthreads=8, dd_local=False, node_wts=None):
        nods, conn = _f(nods), _f(conn, np.int64)
        dd, pts = _f(dd), _f(pts)
        w = _wts(node_wts)
        n_pts = pts.shape[1]
        out = np.zeros(6 * n_pts)
        s = np.ascontiguousarray(s_inf, dtype=np.float64)
        self._stress(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], conn.shape[1], G, nu,
                     _p(dd), _p(pts), n_pts, _p(s), _pw(w), int(dd_local), nthreads, _p(out))
        return out.reshape(n_pts, 6, order="F")

    def getrf(self, A, nthreads=8):
        LU = _f(A).copy(order="F")
        n = LU.shape[0]
        ipiv = np.zeros(n, dtype=np.int32)
        info = self._getrf(n, _p(LU), n, _p(ipiv), nthreads)
        return LU, ipiv, int(info)

    def getrs(self, LU, ipiv, B):
        n = LU.shape[0]
        X = _f(B).copy(order="F")
        nrhs = 1 if X.ndim == 1 else X.shape[1]
        self._getrs(n, nrhs, _p(_f(LU)), n, _p(np.ascontiguousarray(ipiv, dtype=np.int32)), _p(X), n)
        return X

    def sf_integral(self, tau, sfm):
        t, s = _sub_pack(tau, sfm)
        out = np.zeros(6)
        self._sfi(_p(t), _p(s), _p(out))
        return out

    def vc_border(self, nods, conn, dof, dd_local=False, node_wts=None):
        """(row, col) of the volume-control border (n_dof each)."""
        nods, conn = _f(nods), _f(conn, np.int64)
        w = _wts(node_wts)
        dof_f = _f(dof, np.int64)
        n_dof = int(dof.max()) + 1
        row, col = np.zeros(n_dof), np.zeros(n_dof)
        self._vcb(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], conn.shape[1], _p(dof_f),
                  _pw(w), int(dd_local), _p(row), _p(col))
        return row, col

    def mod_system_vc(self, A, orig_dof, dof, used_ndof, delta_t, delta_v):
        """(M, rhs) of the truncated volume-control system; A is the
        (orig_ndof+1)^2 volume-control matrix."""
        A = _f(A)
        od, ud = _f(orig_dof, np.int64), _f(dof, np.int64)
        n_el, ndpe = ud.shape
        nu = used_ndof + 1
        M, rhs = np.zeros(nu * nu), np.zeros(nu)
        dt = np.ascontiguousarray(delta_t, dtype=np.float64)
        rc = self._mvc(_p(A), A.shape[0], A.shape[0] - 1, _p(od), _p(ud), n_el, ndpe, used_ndof, _p(dt), dt.size,
                       float(delta_v), _p(M), _p(rhs))
        if rc != 0:
            raise ValueError("invalid truncation")
        return M.reshape(nu, nu, order="F"), rhs

    def vc_cf_iteration(self, nods, conn, orig_dof, A, mat_id, cmid, par, mr_open, mr_slip, s_inf, inj,
                        new_vol, dd, pp, dof, beta=0.125, dd_local=False):
        """One step of vc_cf_iteration.  dd (6 n_el x 3), pp (6 n_el) and dof
        (n_el x 18, the current handle) are not modified; returns a dict with
        the new dd, pp, dof, n_dof, friction_coef, slip_cohesion,
        open_cohesion, delta_trac, incr, pressure, residual and rc."""
        nods, conn = _f(nods), _f(conn, np.int64)
        A = _f(A)
        od = _f(orig_dof, np.int64)
        n_el = od.shape[0]
        n0 = A.shape[0] - 1
        dd = np.array(_f(dd))
        pp = np.array(pp, dtype=np.float64)
        dof = np.array(_f(dof, np.int64), order="F")
        mid = np.ascontiguousarray(mat_id, dtype=np.int32)
        cm = _f(cmid, np.int32)
        pr = np.ascontiguousarray(par, dtype=np.float64).reshape(-1, 6)
        mo = np.ascontiguousarray(mr_open, dtype=np.float64)
        ms = np.ascontiguousarray(mr_slip, dtype=np.float64)
        si = np.ascontiguousarray(s_inf, dtype=np.float64)
        n_nod = 6 * n_el
        fc, sc, oc = np.zeros(n_nod), np.zeros(n_nod), np.zeros(n_nod)
        dtr, incr = np.zeros(n0), np.zeros(n0 + 1)
        nd, p_new, res = _i64(0), _d(0.0), _d(0.0)
        rc = self._cfi(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], n_el, float(beta),
                       int(dd_local), _p(od), n0, _p(A), A.shape[0], _p(mid), _p(cm), pr.shape[0], _p(pr), _p(mo),
                       _p(ms), _p(si), int(inj[0]), int(inj[1]), float(new_vol), _p(dd), _p(pp), _p(dof),
                       ctypes.byref(nd), _p(fc), _p(sc), _p(oc), _p(dtr), _p(incr), ctypes.byref(p_new),
                       ctypes.byref(res))
        if rc < 0:
            raise ValueError("invalid vc_cf_iteration input")
        return dict(dd=dd, pp=pp, dof=dof, n_dof=nd.value, friction_coef=fc, slip_cohesion=sc, open_cohesion=oc,
                    delta_trac=dtr, incr=incr[:nd.value + 1], pressure=p_new.value, residual=res.value, rc=rc)

    def sneddon_dd(self, G, nu, a, p, xyz):
        dd = np.zeros(3)
        self._sned(G, nu, a, p, _p(np.ascontiguousarray(xyz, dtype=np.float64)), _p(dd))
        return dd


class Reference(_FrictionCohesion):
    """The reference's own kernel/assembly/stress code (oracle/_ref)."""

    def __init__(self, path=REF):
        L = _Lib(path, "ref_")
        self._sub = L.fn("local_submatrix", _int, [_d, _d, _d, _d, _d, _dp, _dp, _dp])
        self._geom = L.fn("el_geom", _int, [_dp, _d, _dp, _dp, _dp, _dp])
        self._geomnu = L.fn("el_geom_nonuniform", _int, [_dp, _dp, _d, _dp, _dp, _dp])
        self._mfc = L.fn("match_f_c", _int, [_i64, _dp, _i64, _dp, _dp, _i64, _dp, _dp, _dp, _dp, _dp, _dp])
        self._mrd = L.fn("update_mrd", _int, [_i64, _dp, _i64, _dp, _i64, _dp, _dp, _dp])
        self._dof = L.fn("dof_h", _i64, [_dp, _i64, _i64, _dp, _i64, _i64, _int, _dp])
        self._asm = L.fn("assemble", _i64, [_dp, _i64, _i64, _dp, _i64, _i64, _d, _d, _d, _int, _int, _int, _dp])
        self._rhs = L.fn("rhs", _i64, [_dp, _i64, _i64, _dp, _i64, _i64, _int, _dp, _dp])
        self._stress = L.fn("stress", _int, [_dp, _i64, _i64, _dp, _i64, _i64, _d, _d, _dp, _dp, _i64, _dp,
                                             _int, _int, _dp])
        self._sfi = L.fn("sf_integral", _int, [_dp, _dp, _dp])
        self._mvcm = L.fn("matrix_vc", _i64, [_dp, _i64, _i64, _dp, _i64, _i64, _d, _d, _d, _int, _int, _int, _dp])
        self._mvc = L.fn("mod_system_vc", _int, [_dp, _i64, _dp, _dp, _i64, _i64, _dp, _i64, _d, _int, _dp, _dp,
                                                  _dp])

    def sf_integral(self, tau, sfm):
        t, s = _sub_pack(tau, sfm)
        out = np.zeros(6)
        self._sfi(_p(t), _p(s), _p(out))
        return out

    def matrix_vc(self, nods, conn, G, nu, beta=0.125, tip_type=0, nthreads=8, dd_local=False):
        """The reference's volume-control matrix ((n_dof+1)^2, corner 0)."""
        if tip_type == 2:
            raise ValueError("the reference's tip_type = 2 numbering writes past n_dof")
        nods, conn = _f(nods), _f(conn, np.int64)
        n_dof, dof = self.dof_h(nods, conn, tip_type)
        A = np.zeros((n_dof + 1) ** 2)
        self._mvcm(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], conn.shape[1], G, nu, beta,
                   tip_type, int(dd_local), nthreads, _p(A))
        return A.reshape(n_dof + 1, n_dof + 1, order="F"), dof

    def mod_system_vc(self, A, orig_dof, dof, used_ndof, delta_t, delta_v, nthreads=1):
        """The reference's truncation; defined only when the original handle
        has no fixed DoF (its border index is n_el * 18)."""
        A = _f(A)
        od, ud = _f(orig_dof, np.int64), _f(dof, np.int64)
        nu = used_ndof + 1
        M, rhs = np.zeros(nu * nu), np.zeros(nu)
        dt = np.ascontiguousarray(delta_t, dtype=np.float64)
        sec = np.zeros(1)
        self._mvc(_p(A), A.shape[0] - 1, _p(od), _p(ud), ud.shape[0], used_ndof, _p(dt), dt.size, float(delta_v),
                  nthreads, _p(M), _p(rhs), _p(sec))
        self.last_seconds = float(sec[0])
        return M.reshape(nu, nu, order="F"), rhs

    def el_geom(self, vert, beta=0.125):
        v = _f(vert)
        sfm, rt, tau, cp = np.zeros(72), np.zeros(9), np.zeros(6), np.zeros(18)
        self._geom(_p(v), beta, _p(sfm), _p(rt), _p(tau), _p(cp))
        return _geom_unpack(sfm, rt, tau, cp)

    def el_geom_nonuniform(self, vert, wts, beta=0.125):
        """(sfm, rt, cp) of make_el_sfm_nonuniform / el_cp_nonuniform."""
        v, w = _f(vert), _wts(wts)
        sfm, rt, cp = np.zeros(72), np.zeros(9), np.zeros(18)
        self._geomnu(_p(v), _p(w), beta, _p(sfm), _p(rt), _p(cp))
        return ((sfm[0::2] + 1j * sfm[1::2]).reshape(6, 6, order="F"), rt.reshape(3, 3, order="F"),
                cp.reshape(3, 6, order="F"))

    def local_submatrix(self, G, nu, h, z, tau, sfm):
        t, s = _sub_pack(tau, sfm)
        out = np.zeros(108)
        self._sub(G, nu, h, z.real, z.imag, _p(t), _p(s), _p(out))
        return out.reshape(6, 18, order="F")

    def dof_h(self, nods, conn, tip_type):
        nods, conn = _f(nods), _f(conn, np.int64)
        n_el = conn.shape[1]
        dof = np.zeros(18 * n_el, dtype=np.int64)
        n = self._dof(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], n_el, tip_type, _p(dof))
        return int(n), dof.reshape(n_el, 18, order="F")

    def assemble(self, nods, conn, G, nu, beta=0.125, tip_type=0, nthreads=8, dd_local=False):
        if tip_type == 2:
            raise ValueError("the reference's tip_type = 2 numbering writes past n_dof")
        nods, conn = _f(nods), _f(conn, np.int64)
        n_dof, dof = self.dof_h(nods, conn, tip_type)
        A = np.zeros(n_dof * n_dof)
        self._asm(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], conn.shape[1], G, nu, beta,
                  tip_type, int(dd_local), nthreads, _p(A))
        return A.reshape(n_dof, n_dof, order="F"), dof

    def rhs(self, nods, conn, tip_type, s_inf):
        nods, conn = _f(nods), _f(conn, np.int64)
        n_dof, _ = self.dof_h(nods, conn, tip_type)
        r = np.zeros(n_dof)
        s = np.ascontiguousarray(s_inf, dtype=np.float64)
        self._rhs(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], conn.shape[1], tip_type,
                  _p(s), _p(r))
        return r

    def stress(self, nods, conn, G, nu, dd, pts, s_inf, nthreads=1, dd_local=False):
        """The reference accumulates stresses from several OpenMP threads into
        shared memory without synchronisation (system_assembly.cpp:415-417):
        values are only reliable with nthreads = 1; more threads are for
        timing."""
        nods, conn = _f(nods), _f(conn, np.int64)
        dd, pts = _f(dd), _f(pts)
        n_pts = pts.shape[1]
        out = np.zeros(6 * n_pts)
        s = np.ascontiguousarray(s_inf, dtype=np.float64)
        self._stress(_p(nods), nods.shape[0], nods.shape[1], _p(conn), conn.shape[0], conn.shape[1], G, nu,
                     _p(dd), _p(pts), n_pts, _p(s), int(dd_local), nthreads, _p(out))
        return out.reshape(n_pts, 6, order="F")


def oracle():
    return Oracle()


def reference_or_none():
    try:
        return Reference()
    except OSError:
        return None


def golden(name):
    return os.path.join(GOLDEN, name)
