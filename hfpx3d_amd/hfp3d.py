"""Python mirror of the reference's hot-path call surface, backed by the HIP
library through its C ABI.

Names, argument meaning and error behaviour follow the reference's namespace
``hfp3d`` so that code written against it reads the same:

  Mesh_Geom_T, DoF_Handle_T, Mesh_Data_T        src/Core/surface_mesh_utilities.h:25-104
  Properties_T, Load_T, Num_Param_T             src/Core/model_parameters.h:34-94
  SAE_T                                         src/Solvers/system_assembly.h:26-30
  make_dof_h_crack                              src/Core/surface_mesh_utilities.cpp:28-112
  make_3dbem_matrix_s                           src/Solvers/system_assembly.cpp:30-219
  add_s_inf_to_3dbem_rhs                        src/Solvers/system_assembly.cpp:222-274
  LU (factor + solve)                           il/linear_algebra/dense/factorization/LU.h:46-138
  write_dd_vector_to_md                         src/Core/surface_mesh_utilities.cpp:256-328
  make_3dbem_stress_f_s                         src/Solvers/system_assembly.cpp:279-455
  make_local_3dbem_submatrix                    src/Elasticity/elasticity_kernel_integration.cpp:34-205

Matrices and right-hand sides live on the GPU as torch tensors in
column-major layout (a logical n x n matrix is the transpose view of a
row-major (n, lda) buffer).  Meshes, DoF handles and DD arrays are host
numpy arrays, as in the reference.  Preconditions that the reference checks
with IL_EXPECT_FAST raise ValueError; numerical failure (a singular matrix)
is reported through a Status object like il::Status.
"""
from dataclasses import dataclass, field
import atexit
import ctypes
import sys

import numpy as np

from . import _lib

# --------------------------------------------------------------------------
# containers
# --------------------------------------------------------------------------


@dataclass
class Mesh_Geom_T:
    nods: np.ndarray = None   # (3 or 6) x n_nodes, float64
    conn: np.ndarray = None   # (3 or 5) x n_el, int64, rows 0..2 0-based vertex ids
    # Optional per-node vertex weights (n_nodes, > 0) of a non-uniform edge
    # partition: the edge node between vertices a and b sits at
    # (w_a x_a + w_b x_b) / (w_a + w_b) (make_el_sfm_nonuniform /
    # el_cp_nonuniform, element_utilities.cpp:240-322, 397-431).  None: the
    # uniform element (edge nodes mid-edge).
    vert_wts: np.ndarray = None


@dataclass
class DoF_Handle_T:
    n_dof: int = 0
    dof_h: np.ndarray = None  # n_el x 18, int64, -1 = fixed


@dataclass
class Num_Param_T:
    beta: float = 0.125
    tip_type: int = 1
    is_dd_local: bool = False


@dataclass
class F_C_Param_T:
    """Friction / cohesion parameters of one surface material
    (src/Core/model_parameters.h:22-30)."""
    cr_open: float = 1.0    # critical opening
    peak_ts: float = 0.0    # peak tensile stress
    cr_slip: float = 1.0    # critical slip
    peak_sc: float = 0.0    # peak shear cohesion
    peak_sf: float = 0.0    # peak friction coefficient
    res_sf: float = 0.0     # residual friction coefficient


@dataclass
class Properties_T:
    n_solid: int = 1
    shear_m: list = field(default_factory=lambda: [1.0])
    poiss_r: list = field(default_factory=lambda: [0.0])
    n_surface: int = 1
    # contact model ids per surface material: column 0 cohesion (0 box,
    # 1 linear softening), column 1 friction (0 constant, 1 linear slip
    # weakening); n_surface x 2
    c_model_id: np.ndarray = None
    f_c_param: list = field(default_factory=list)   # F_C_Param_T per surface material


@dataclass
class Frac_State_T:
    """Node-wise crack-surface state (src/Core/surface_mesh_utilities.h:51-62)."""
    mr_open: np.ndarray = None        # largest relative opening reached
    mr_slip: np.ndarray = None        # largest relative slip reached
    friction_coef: np.ndarray = None
    slip_cohesion: np.ndarray = None
    open_cohesion: np.ndarray = None

    @classmethod
    def zeros(cls, n_nod):
        return cls(*(np.zeros(n_nod) for _ in range(5)))


@dataclass
class Load_T:
    s_inf: np.ndarray = field(default_factory=lambda: np.zeros(6))
    # injection locations: rows (element, local node); the first is the
    # injection point of the volume-control iteration (model_parameters.h:64)
    inj_loc: np.ndarray = None


@dataclass
class Mesh_Data_T:
    mesh: Mesh_Geom_T = None
    dof_h_dd: DoF_Handle_T = field(default_factory=DoF_Handle_T)
    dd: np.ndarray = None     # (6 n_el) x 3
    dof_h_pp: DoF_Handle_T = field(default_factory=DoF_Handle_T)  # n_el x 6 pressure DoF, -1 = none
    pp: np.ndarray = None     # fluid pressure per node (6 n_el)
    ae_set: list = field(default_factory=list)  # elements with a DD DoF ("active")
    fe_set: list = field(default_factory=list)  # elements with a pressure DoF ("filled")
    mat_id: np.ndarray = None  # surface material id per node (6 n_el)
    frac_state: "Frac_State_T" = None


@dataclass
class SAE_T:
    n_dof: int = 0
    matrix: object = None     # torch tensor, n_dof x n_dof, column-major, device
    rhs_v: object = None      # torch tensor, n_dof (or n_dof x nrhs), device
    matrix32: object = None   # optional FP32 rounding of matrix (same shape)


class Status:
    """Minimal il::Status: ok() / error / info (rank for a singular matrix)."""

    def __init__(self):
        self.error = None
        self.info = 0

    def ok(self):
        return self.error is None

    def abortOnError(self):
        if self.error is not None:
            raise RuntimeError(f"{self.error} (info = {self.info})")


# --------------------------------------------------------------------------
# context (one per device, on a dedicated default-priority torch stream)
# --------------------------------------------------------------------------

# Set once the contexts have been closed at interpreter exit: objects that
# are collected after that point release nothing through HIP (the process's
# device memory goes with it).
_shut_down = False


def _hip_calls_allowed():
    return not _shut_down and not sys.is_finalizing()


class Context:
    def __init__(self, device=0):
        import torch
        _register_exit_hook()
        self.torch = torch
        self.device = torch.device("cuda", device)
        self.lib = _lib.load()
        self.stream = torch.cuda.Stream(device=self.device)
        self.autotune = None  # table of knob settings by (n-range, precision), see autotune.py
        h = ctypes.c_void_p()
        _lib.check(self.lib.hfpx_create(ctypes.byref(h), device, ctypes.c_void_p(self.stream.cuda_stream)))
        self.handle = h

    def enter(self):
        """Order this context's stream after the caller's current stream."""
        self.stream.wait_stream(self.torch.cuda.current_stream(self.device))

    def leave(self):
        """Order the caller's current stream after this context's work."""
        self.torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def check(self, rc):
        _lib.check(rc, self.handle)

    def close(self):
        """Finish the context's work and free its device memory.  The stream
        is torch's, so it is synchronised here, while torch is alive; the
        library only frees what it allocated."""
        if self.handle:
            if _hip_calls_allowed():
                self.stream.synchronize()
                self.lib.hfpx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}
_exit_hook = False


def close_all():
    """Close every context of get_context().  Runs at interpreter exit
    (registered after torch is imported, so it runs before torch's own exit
    handlers), and may be called earlier, e.g. at the end of a benchmark."""
    for ctx in list(_contexts.values()):
        try:
            ctx.close()
        except Exception:
            pass
    _contexts.clear()


def _at_exit():
    global _shut_down
    close_all()
    _shut_down = True


def _register_exit_hook():
    global _exit_hook
    if not _exit_hook:
        atexit.register(_at_exit)
        _exit_hook = True


def get_context(device=0):
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    return ctx


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DeviceMesh:
    """A mesh (plus DoF map and collocation offset) resident on the GPU."""

    def __init__(self, mesh, dof_hndl=None, beta=0.125, ctx=None, dd_local=False):
        self.ctx = ctx or get_context()
        nods = np.asfortranarray(mesh.nods, dtype=np.float64)
        conn = np.asfortranarray(mesh.conn, dtype=np.int64)
        if conn.shape[0] < 3 or conn.shape[1] < 1 or nods.shape[0] < 3 or nods.shape[1] < 3:
            raise ValueError("mesh needs >= 3 rows of coordinates/connectivity, >= 3 nodes, >= 1 element")
        self.n_el = conn.shape[1]
        dof_ptr, n_dof = None, 18 * self.n_el
        if dof_hndl is not None and dof_hndl.dof_h is not None and dof_hndl.n_dof > 0:
            dh = np.asfortranarray(dof_hndl.dof_h, dtype=np.int64)
            if dh.shape != (self.n_el, 18):
                raise ValueError("dof_h must be n_el x 18")
            self._dh = dh
            dof_ptr, n_dof = _np_ptr(dh), int(dof_hndl.n_dof)
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.hfpx_mesh_create(
            self.ctx.handle, _np_ptr(nods), nods.shape[0], nods.shape[1], _np_ptr(conn), conn.shape[0],
            self.n_el, dof_ptr, n_dof, float(beta), ctypes.byref(h)))
        self.handle = h
        self.n_dof = int(self.ctx.lib.hfpx_mesh_n_dof(h))
        self.set_dd_local(dd_local)
        self.set_vertex_weights(getattr(mesh, "vert_wts", None))

    def set_vertex_weights(self, wts):
        """Per-node vertex weights of a non-uniform edge partition (n_nodes,
        all > 0), or None for the uniform element."""
        if wts is None:
            self.ctx.check(self.ctx.lib.hfpx_mesh_set_vertex_weights(self.ctx.handle, self.handle, None, 0))
        else:
            w = np.ascontiguousarray(wts, dtype=np.float64).reshape(-1)
            self.ctx.check(self.ctx.lib.hfpx_mesh_set_vertex_weights(self.ctx.handle, self.handle, _np_ptr(w),
                                                                     w.size))
        self.vert_wts = wts

    def set_dd_local(self, dd_local):
        """DD in each element's own frame (Num_Param_T::is_dd_local)."""
        self.ctx.check(self.ctx.lib.hfpx_mesh_set_dd_local(self.handle, int(bool(dd_local))))
        self.dd_local = bool(dd_local)

    def close(self):
        if getattr(self, "handle", None):
            if _hip_calls_allowed():
                self.ctx.lib.hfpx_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def alloc_colmajor(n_rows, n_cols, dtype, ctx, lda=None):
    """Device buffer for an n_rows x n_cols column-major matrix (lda >= n_rows);
    returns (logical view with strides (1, lda), lda)."""
    torch = ctx.torch
    lda = lda or n_rows
    with torch.cuda.stream(ctx.stream):
        buf = torch.empty((n_cols, lda), dtype=dtype, device=ctx.device)
    return buf.t()[:n_rows], lda


# --------------------------------------------------------------------------
# DoF handle (host)
# --------------------------------------------------------------------------

def make_dof_h_crack(mesh, ap_order=2, tip_type=1):
    """Per-element DoF map with tip DoF fixed (-1) and the rest numbered
    consecutively.  A vertex is at the tip when node rows 3 and 4 are both 1
    and the element's surface ids (conn rows 3, 4) are equal; tip_type 2 also
    fixes mid-edge nodes between two tip vertices.  Unlike the reference
    (surface_mesh_utilities.cpp:85-101), every free DoF is renumbered
    consecutively for tip_type 2 as well: the reference leaves mid-edge DoF
    of non-tip edges un-renumbered there, which puts indices past n_dof."""
    if ap_order != 2:
        raise ValueError("only quadratic (ap_order = 2) elements are supported")
    nods, conn = np.asarray(mesh.nods), np.asarray(mesh.conn)
    n_el = conn.shape[1]
    ext = nods.shape[0] >= 5 and conn.shape[0] >= 5
    dof = np.zeros((n_el, 18), dtype=np.int64)
    n_dof, dec = 18 * n_el, 0
    for e in range(n_el):
        tip = [False, False, False]
        if ext and conn[3, e] == conn[4, e]:
            for v in range(3):
                n = conn[v, e]
                tip[v] = nods[3, n] == 1 and nods[4, n] == 1
        for v in range(6):
            fixed = (tip_type >= 1 and v < 3 and tip[v]) or \
                    (tip_type == 2 and v >= 3 and tip[(v - 2) % 3] and tip[(v - 1) % 3])
            for l in range(3):
                ld = 3 * v + l
                if fixed:
                    dof[e, ld] = -1
                    n_dof -= 1
                    dec += 1
                else:
                    dof[e, ld] = 18 * e + ld - dec
    return DoF_Handle_T(n_dof=n_dof, dof_h=np.asfortranarray(dof))


# --------------------------------------------------------------------------
# assembly and right-hand side
# --------------------------------------------------------------------------

def make_3dbem_matrix_s(shear_m, poiss_r, mesh, n_par, dof_hndl=None, ctx=None, dtype=None,
                        device_mesh=None):
    """Dense n_dof x n_dof traction influence matrix on the GPU (FP64 unless
    dtype=torch.float32).  If dof_hndl is empty it is built with
    make_dof_h_crack(mesh, 2, n_par.tip_type) and filled in, like the
    reference's io argument."""
    ctx = ctx or get_context()
    torch = ctx.torch
    dtype = dtype or torch.float64
    if dof_hndl is not None and (dof_hndl.n_dof == 0 or dof_hndl.dof_h is None):
        d = make_dof_h_crack(mesh, 2, n_par.tip_type)
        dof_hndl.n_dof, dof_hndl.dof_h = d.n_dof, d.dof_h
    dm = device_mesh or DeviceMesh(mesh, dof_hndl, n_par.beta, ctx)
    dm.set_dd_local(n_par.is_dd_local)
    A, lda = alloc_colmajor(dm.n_dof, dm.n_dof, dtype, ctx)
    ctx.enter()
    if dtype == torch.float64:
        ctx.check(ctx.lib.hfpx_assemble(ctx.handle, dm.handle, float(shear_m), float(poiss_r), _ptr(A), lda))
    elif dtype == torch.float32:
        ctx.check(ctx.lib.hfpx_assemble_mixed(ctx.handle, dm.handle, float(shear_m), float(poiss_r),
                                              None, 0, _ptr(A), lda))
    else:
        raise ValueError("dtype must be float64 or float32")
    ctx.leave()
    return A


def add_s_inf_to_3dbem_rhs(mesh_data, load, sae, ctx=None, device_mesh=None):
    """sae.rhs_v -= traction of the far-field stress load.s_inf on every free
    DoF.  load.s_inf may be 6 or 6 x nrhs (order xx yy zz xy xz yz)."""
    ctx = ctx or get_context()
    torch = ctx.torch
    s = np.asfortranarray(np.asarray(load.s_inf, dtype=np.float64).reshape(6, -1))
    nrhs = s.shape[1]
    dm = device_mesh or DeviceMesh(mesh_data.mesh, mesh_data.dof_h_dd, 0.125, ctx)
    n = dm.n_dof
    rhs = sae.rhs_v
    if rhs is None or rhs.shape[0] != n:
        if nrhs == 1:
            with torch.cuda.stream(ctx.stream):
                rhs = torch.zeros(n, dtype=torch.float64, device=ctx.device)
        else:
            rhs, _ = alloc_colmajor(n, nrhs, torch.float64, ctx)
            with torch.cuda.stream(ctx.stream):
                rhs.zero_()
    ldb = n if rhs.dim() == 1 else rhs.stride(1)
    ctx.enter()
    ctx.check(ctx.lib.hfpx_rhs_sinf(ctx.handle, dm.handle, _np_ptr(s), nrhs, _ptr(rhs), ldb))
    ctx.leave()
    sae.rhs_v = rhs
    sae.n_dof = n
    return rhs


# --------------------------------------------------------------------------
# volume control
# --------------------------------------------------------------------------

def make_3dbem_matrix_vc(shear_m, poiss_r, mesh, n_par, dof_hndl=None, ctx=None, dtype=None, device_mesh=None,
                         with_f32=False):
    """Volume-control matrix ((n_dof+1) x (n_dof+1), device, column-major):
    the influence matrix of make_3dbem_matrix_s bordered by the opening-volume
    row (n_dof) and the traction-per-pressure column (n_dof), corner 0.
    Mirrors hfp3d::make_3dbem_matrix_vc (system_assembly.cpp:458-557),
    including the io fill-in of an empty dof_hndl.  dtype=torch.float32 gives
    the FP32 rounding; with_f32=True returns (FP64 matrix, FP32 matrix) from
    one assembly pass."""
    ctx = ctx or get_context()
    torch = ctx.torch
    dtype = dtype or torch.float64
    if dof_hndl is not None and (dof_hndl.n_dof == 0 or dof_hndl.dof_h is None):
        d = make_dof_h_crack(mesh, 2, n_par.tip_type)
        dof_hndl.n_dof, dof_hndl.dof_h = d.n_dof, d.dof_h
    dm = device_mesh or DeviceMesh(mesh, dof_hndl, n_par.beta, ctx)
    dm.set_dd_local(n_par.is_dd_local)
    n1 = dm.n_dof + 1
    A64 = A32 = None
    if with_f32 or dtype == torch.float64:
        A64, _ = alloc_colmajor(n1, n1, torch.float64, ctx)
    if with_f32 or dtype == torch.float32:
        A32, _ = alloc_colmajor(n1, n1, torch.float32, ctx)
    if A64 is None and A32 is None:
        raise ValueError("dtype must be float64 or float32")
    ctx.enter()
    ctx.check(ctx.lib.hfpx_assemble_vc(ctx.handle, dm.handle, float(shear_m), float(poiss_r),
                                       None if A64 is None else _ptr(A64), n1,
                                       None if A32 is None else _ptr(A32), n1))
    ctx.leave()
    if with_f32:
        return A64, A32
    return A64 if A64 is not None else A32


def mod_3dbem_system_vc(orig_matrix, orig_dof_hndl, dof_hndl, delta_t, delta_v, ctx=None, with_f32=False):
    """Truncated volume-control system gathered on the device from the
    resident (orig_ndof+1)^2 matrix orig_matrix: SAE_T with n_dof = used DoF,
    matrix (used+1)^2 FP64 (and matrix32, its FP32 rounding, when with_f32),
    rhs_v (used+1) = [delta_t on the used DoF, delta_v].  delta_t is indexed
    per slot, per original DoF or per used DoF by its length (n_el * 18,
    orig n_dof or used n_dof).  Mirrors hfp3d::mod_3dbem_system_vc
    (system_assembly.cpp:560-629); the border is read at row/column
    orig n_dof (the reference's n_el * 18 when no original DoF is fixed)."""
    ctx = ctx or get_context()
    torch = ctx.torch
    n_o, lda = _check_colmajor(orig_matrix)
    if orig_matrix.dtype != torch.float64:
        raise ValueError("orig_matrix must be the FP64 volume-control matrix")
    if n_o != orig_dof_hndl.n_dof + 1:
        raise ValueError(f"orig_matrix is {n_o} x {n_o}, expected orig n_dof + 1 = {orig_dof_hndl.n_dof + 1}")
    od = np.asfortranarray(orig_dof_hndl.dof_h, dtype=np.int64)
    ud = np.asfortranarray(dof_hndl.dof_h, dtype=np.int64)
    if od.shape != ud.shape:
        raise ValueError("DoF handles of different shapes")
    n_el, ndpe = ud.shape
    used = int(dof_hndl.n_dof)
    nu = used + 1
    dt = np.ascontiguousarray(np.asarray(delta_t, dtype=np.float64).reshape(-1))
    M, ldm = alloc_colmajor(nu, nu, torch.float64, ctx)
    M32 = None
    if with_f32:
        M32, _ = alloc_colmajor(nu, nu, torch.float32, ctx)
    with torch.cuda.stream(ctx.stream):
        rhs = torch.empty(nu, dtype=torch.float64, device=ctx.device)
    ctx.enter()
    ctx.check(ctx.lib.hfpx_vc_truncate(ctx.handle, _ptr(orig_matrix), lda, int(orig_dof_hndl.n_dof), _np_ptr(od),
                                       _np_ptr(ud), n_el, ndpe, used, _np_ptr(dt), dt.size, float(delta_v), _ptr(M),
                                       ldm, None if M32 is None else _ptr(M32), nu, _ptr(rhs)))
    ctx.leave()
    return SAE_T(n_dof=used, matrix=M, rhs_v=rhs, matrix32=M32)


def write_dd_vector_to_md(sol_v, dof_h_dd, include_p=False, dof_h_pp=None, m_data=None):
    """Solution vector -> nodal DD array ((6 n_el) x 3; fixed DoF stay 0),
    returned.  With include_p, the pressure unknowns that follow the
    dof_h_dd.n_dof DD unknowns go to nodal pressures pp[6 el + node] =
    sol_v[n_dd + dof_h_pp.dof_h[el, node]] (dof_h_pp n_el x 6, -1 = none).
    With m_data (Mesh_Data_T), its dd (kept where fixed, created when
    empty), pp, ae_set and fe_set (elements with a DD / pressure DoF) are
    written as the reference's io argument.  Mirrors
    hfp3d::write_dd_vector_to_md (src/Core/surface_mesh_utilities.cpp:
    256-328); pp holds one value per node (the reference sizes it per
    element and indexes it per node).  Size mismatches raise ValueError
    (the reference's IL_EXPECT_FAST)."""
    sol = np.asarray(sol_v.cpu() if hasattr(sol_v, "cpu") else sol_v, dtype=np.float64).reshape(-1)
    dh = np.asarray(dof_h_dd.dof_h)
    n_el = dh.shape[0]
    if dh.shape != (n_el, 18):
        raise ValueError("dof_h_dd must be n_el x 18")
    n_dd = int(dof_h_dd.n_dof)
    n_pp = int(dof_h_pp.n_dof) if include_p else 0
    if sol.size != n_dd + n_pp:
        raise ValueError(f"sol_v has {sol.size} entries, expected n_dof(dd) + n_dof(pp) = {n_dd + n_pp}")
    if m_data is not None and m_data.dd is not None and np.asarray(m_data.dd).size:
        dd = np.array(m_data.dd, dtype=np.float64)
        if dd.shape != (6 * n_el, 3):
            raise ValueError("m_data.dd must be (6 n_el) x 3")
    else:
        dd = np.zeros((6 * n_el, 3))
    d = dh.reshape(n_el, 6, 3)
    if d.max(initial=-1) >= n_dd:
        raise ValueError("DD DoF out of range")
    mask = d >= 0
    v = dd.reshape(n_el, 6, 3)
    v[mask] = sol[d[mask]]
    active = mask.any(axis=(1, 2))
    pp, filled = None, np.zeros(n_el, dtype=bool)
    if include_p:
        ph = np.asarray(dof_h_pp.dof_h)
        if ph.shape != (n_el, 6):
            raise ValueError("dof_h_pp must be n_el x 6")
        if ph.max(initial=-1) >= n_pp:
            raise ValueError("pressure DoF out of range")
        pp = None if m_data is None or getattr(m_data, "pp", None) is None else np.array(m_data.pp, dtype=np.float64)
        if pp is None or pp.size == 0:
            pp = np.zeros(6 * n_el)
        if pp.shape != (6 * n_el,):
            raise ValueError("m_data.pp must hold 6 n_el values")
        pm = ph >= 0
        pp.reshape(n_el, 6)[pm] = sol[n_dd + ph[pm]]
        filled = pm.any(axis=1)
    if m_data is not None:
        m_data.dd = dd
        if include_p:
            m_data.pp = pp
        m_data.ae_set = list(getattr(m_data, "ae_set", None) or []) + [int(e) for e in np.flatnonzero(active)]
        m_data.fe_set = list(getattr(m_data, "fe_set", None) or []) + [int(e) for e in np.flatnonzero(filled)]
    return dd


# --------------------------------------------------------------------------
# the element integral
# --------------------------------------------------------------------------

def make_local_3dbem_submatrix(kernel_id, shear_m, poiss_r, h, z, tau, sfm, ctx=None):
    """6 x 18 stress influence (rows S11 S22 S33 S12 S13 S23 in the element
    frame, column 3 node + DD component) of the element with vertices tau
    (3 complex) and shape-function coefficients sfm (6 x 6 complex, row =
    node, column = monomial 1, tau, conj(tau), tau^2, conj(tau)^2,
    tau conj(tau)) at the point (h, z) of the element's frame, on the GPU.
    Batched: h (n,), z (n,), tau (n, 3), sfm (n, 6, 6) give (n, 6, 18).
    kernel_id 1 (hypersingular) only.  Mirrors
    hfp3d::make_local_3dbem_submatrix
    (src/Elasticity/elasticity_kernel_integration.cpp:34-205)."""
    ctx = ctx or get_context()
    torch = ctx.torch
    hh = np.atleast_1d(np.asarray(h, dtype=np.float64))
    single = np.ndim(h) == 0
    n = hh.size
    zz = np.asarray(z, dtype=np.complex128).reshape(n)
    tt = np.asarray(tau, dtype=np.complex128).reshape(n, 3)
    ss = np.asarray(sfm, dtype=np.complex128).reshape(n, 6, 6)
    zv = np.stack([zz.real, zz.imag], axis=1).reshape(-1)
    tv = np.stack([tt.real, tt.imag], axis=2).reshape(-1)
    sc = np.transpose(ss, (0, 2, 1))  # column-major 6 x 6 per case
    sv = np.stack([sc.real, sc.imag], axis=3).reshape(-1)
    with torch.cuda.stream(ctx.stream):
        d = [torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device) for a in (hh, zv, tv, sv)]
        out = torch.empty(108 * n, dtype=torch.float64, device=ctx.device)
    ctx.enter()
    ctx.check(ctx.lib.hfpx_local_submatrix(ctx.handle, int(kernel_id), float(shear_m), float(poiss_r), n,
                                           *[_ptr(t) for t in d], _ptr(out)))
    ctx.leave()
    M = out.cpu().numpy().reshape(n, 18, 6).transpose(0, 2, 1)
    return M[0] if single else M


# --------------------------------------------------------------------------
# LU (il::LU semantics: the matrix is taken over, factorised in place)
# --------------------------------------------------------------------------

def _dtype_code(torch, t):
    if t.dtype == torch.float64:
        return _lib.HFPX_F64
    if t.dtype == torch.float32:
        return _lib.HFPX_F32
    raise ValueError("LU supports float64 and float32 matrices")


def _check_colmajor(A):
    n = A.shape[0]
    if A.dim() != 2 or A.shape[1] != n or A.stride(0) != 1 or A.stride(1) < n:
        raise ValueError("expected a square column-major device matrix (strides (1, lda))")
    return n, A.stride(1)


class LU:
    """LU factorisation with partial pivoting on the GPU.

    Mirrors il::LU<il::Array2D<double>> (il/linear_algebra/dense/factorization/
    LU.h:46-138): the constructor factorises the given matrix (which it takes
    over: the storage is overwritten with the factors), reports an exactly
    singular matrix through `status` (error 'MatrixSingular', info = rank,
    LU.h:94-98), and solve() applies the factors to one or many right-hand
    sides (dgetrs)."""

    def __init__(self, A, status=None, ctx=None):
        self.ctx = ctx or get_context()
        torch = self.ctx.torch
        self.n, self.lda = _check_colmajor(A)
        self.A = A
        self.code = _dtype_code(torch, A)
        with torch.cuda.stream(self.ctx.stream):
            self.ipiv = torch.empty(self.n, dtype=torch.int32, device=self.ctx.device)
        if self.ctx.autotune is not None:
            from . import autotune
            autotune.apply(self.ctx.autotune, self.n, "f32" if A.dtype == torch.float32 else "f64", self.ctx)
        info = ctypes.c_int64(0)
        self.ctx.enter()
        rc = self.ctx.lib.hfpx_getrf(self.ctx.handle, self.code, _ptr(A), self.lda, self.n, _ptr(self.ipiv),
                                     ctypes.byref(info))
        self.ctx.leave()
        self.info = int(info.value)
        if rc == _lib.HFPX_ERR_SINGULAR:
            if status is None:
                raise _lib.HfpxError(rc, f"matrix is singular (U({self.info},{self.info}) = 0)")
            status.error = "MatrixSingular"
            status.info = self.info - 1
        else:
            self.ctx.check(rc)

    def solve(self, y):
        """Solve A x = y for a device vector (n) or column-major matrix
        (n x nrhs); returns a new tensor, y is left unchanged."""
        torch = self.ctx.torch
        with torch.cuda.stream(self.ctx.stream):
            if y.dim() == 1:
                x = y.to(dtype=self.A.dtype, device=self.ctx.device).clone()
                nrhs, ldb = 1, self.n
            else:
                x, ldb = alloc_colmajor(self.n, y.shape[1], self.A.dtype, self.ctx)
                x.copy_(y)
                nrhs = y.shape[1]
        self.ctx.enter()
        self.ctx.check(self.ctx.lib.hfpx_getrs(self.ctx.handle, self.code, _ptr(self.A), self.lda,
                                               _ptr(self.ipiv), self.n, _ptr(x), ldb, nrhs))
        self.ctx.leave()
        return x


def gemm_sub(C, A, B, ctx=None):
    """C -= A @ B on MFMA for column-major device matrices (testing aid)."""
    ctx = ctx or get_context()
    torch = ctx.torch
    M, N = C.shape
    K = A.shape[1]
    for t in (C, A, B):
        if t.stride(0) != 1:
            raise ValueError("column-major operands expected")
    ctx.enter()
    ctx.check(ctx.lib.hfpx_gemm_sub(ctx.handle, _dtype_code(torch, C), _ptr(C), C.stride(1), _ptr(A), A.stride(1),
                                    _ptr(B), B.stride(1), M, N, K))
    ctx.leave()
    return C


def tune_set(knob, value, ctx=None):
    """Select a kernel variant on the context (hfpx_tune_set)."""
    ctx = ctx or get_context()
    ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, int(knob), int(value)))


def tune_get(knob, ctx=None):
    ctx = ctx or get_context()
    v = ctypes.c_int(0)
    ctx.check(ctx.lib.hfpx_tune_get(ctx.handle, int(knob), ctypes.byref(v)))
    return v.value


def solve_mixed(A64, B, tol=1e-14, max_outer=10, gmres_m=20, ctx=None, LU32=None):
    """Solve A64 X = B: FP32 LU of A64 (or the given FP32 LU object) as the
    preconditioner of FP64 iterative refinement against A64 (classic steps
    while they contract the backward error 100-fold, then GMRES(gmres_m)).
    Returns (X, report dict: iterations, inner_iterations, backward_error,
    converged, fell_back (the solution came from an FP64 LU of a copy of
    A64), status (HFPX_OK or HFPX_ERR_NOT_CONVERGED)).  An FP32 factor
    with a zero pivot is accepted (the refinement then falls back)."""
    ctx = ctx or get_context()
    torch = ctx.torch
    n, lda = _check_colmajor(A64)
    if LU32 is None:
        A32, _ = alloc_colmajor(n, n, torch.float32, ctx)
        with torch.cuda.stream(ctx.stream):
            A32.copy_(A64)
        LU32 = LU(A32, ctx=ctx)
    B2 = B if B.dim() == 2 else B.reshape(n, 1)
    nrhs = B2.shape[1]
    ldb = B2.stride(1) if nrhs > 1 else n
    X, ldx = alloc_colmajor(n, nrhs, torch.float64, ctx)
    rep = _lib.IRReport()
    ctx.enter()
    rc = ctx.lib.hfpx_solve_ir(ctx.handle, _ptr(A64), lda, _ptr(LU32.A), LU32.lda, _ptr(LU32.ipiv), n, _ptr(B2), ldb,
                               _ptr(X), ldx, nrhs, float(tol), int(max_outer), int(gmres_m), ctypes.byref(rep))
    ctx.leave()
    report = dict(iterations=rep.iterations, inner_iterations=rep.inner_iterations,
                  backward_error=rep.backward_error, converged=bool(rep.converged), fell_back=bool(rep.fell_back),
                  status=rc)
    if rc not in (_lib.HFPX_OK, _lib.HFPX_ERR_NOT_CONVERGED):
        ctx.check(rc)
    return (X if B.dim() == 2 else X[:, 0]), report


# --------------------------------------------------------------------------
# stress postprocessing
# --------------------------------------------------------------------------

def make_3dbem_stress_f_s(shear_m, poiss_r, m_data, load, n_par, m_pts_crd, ctx=None, device_mesh=None):
    """Stresses (n_pts x 6: Sxx Syy Szz Sxy Sxz Syz, tension positive) at the
    points m_pts_crd (3 x n_pts) due to the nodal DD m_data.dd ((6 n_el) x 3,
    global components, or each element's own frame when n_par.is_dd_local)
    plus load.s_inf, like the reference (the far-field stress is added to
    every point).  Returns a numpy array."""
    ctx = ctx or get_context()
    torch = ctx.torch
    pts = np.asarray(m_pts_crd, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[0] != 3:
        raise ValueError("m_pts_crd must be 3 x n_pts")
    n_pts = pts.shape[1]
    dd = np.asarray(m_data.dd, dtype=np.float64)
    n_el = np.asarray(m_data.mesh.conn).shape[1]
    if dd.shape != (6 * n_el, 3):
        raise ValueError("m_data.dd must be (6 n_el) x 3")
    dm = device_mesh or DeviceMesh(m_data.mesh, None, 0.0, ctx)
    dm.set_dd_local(n_par is not None and n_par.is_dd_local)
    with torch.cuda.stream(ctx.stream):
        d_dd = torch.from_numpy(np.ascontiguousarray(dd.T)).to(ctx.device)      # (3, 6 n_el): col-major dd
        d_pts = torch.from_numpy(np.ascontiguousarray(pts.T)).to(ctx.device)    # (n_pts, 3): col-major pts
        out = torch.empty((6, n_pts), dtype=torch.float64, device=ctx.device)   # col-major n_pts x 6
    s_inf = np.ascontiguousarray(np.asarray(load.s_inf, dtype=np.float64).reshape(6))
    ctx.enter()
    ctx.check(ctx.lib.hfpx_stress(ctx.handle, dm.handle, float(shear_m), float(poiss_r), _ptr(d_dd), 6 * n_el,
                                  _ptr(d_pts), n_pts, _np_ptr(s_inf), _ptr(out), n_pts))
    ctx.leave()
    return out.t().cpu().numpy()


# --------------------------------------------------------------------------
# batches of independent cracks (no reference counterpart)
# --------------------------------------------------------------------------

class Batch:
    """Many independent cracks assembled and solved together on the GPU
    (hfpx_batch_*).  `cracks` is a list of dicts with nods (3+ x n_nodes),
    conn (3+ x n_el, 0-based vertex ids in rows 0..2), G, nu and s_inf (6,
    xx yy zz xy xz yz, tension positive); every DoF is free (fix_tip = 0).
    Cracks with n >= mixed_min_n are solved by FP32 LU + FP64 refinement,
    the others by FP64 LU."""

    def __init__(self, cracks, mixed_min_n=1024, beta=0.125, ctx=None):
        self.ctx = ctx or get_context()
        self.lib = self.ctx.lib
        self.n_cracks = len(cracks)
        n_nodes = np.array([c["nods"].shape[1] for c in cracks], dtype=np.int64)
        n_el = np.array([c["conn"].shape[1] for c in cracks], dtype=np.int64)
        nods = np.ascontiguousarray(np.concatenate([np.asarray(c["nods"], dtype=np.float64)[:3].T for c in cracks]))
        conn = np.ascontiguousarray(np.concatenate([np.asarray(c["conn"], dtype=np.int64)[:3].T for c in cracks]))
        G = np.array([c["G"] for c in cracks], dtype=np.float64)
        nu = np.array([c["nu"] for c in cracks], dtype=np.float64)
        sinf = np.ascontiguousarray(np.array([np.asarray(c["s_inf"], dtype=np.float64) for c in cracks]))
        self.n = 18 * n_el
        h = ctypes.c_void_p()
        self.ctx.check(self.lib.hfpx_batch_create(self.ctx.handle, self.n_cracks, _np_ptr(n_nodes), _np_ptr(n_el),
                                                  _np_ptr(nods), _np_ptr(conn), float(beta), _np_ptr(G), _np_ptr(nu),
                                                  _np_ptr(sinf), int(mixed_min_n), ctypes.byref(h)))
        self.handle = h

    def close(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _hip_calls_allowed():
            self.lib.hfpx_batch_destroy(h)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble(self):
        self.ctx.check(self.lib.hfpx_batch_assemble(self.ctx.handle, self.handle))

    def solve(self, n_refine=3, use_graph=True):
        self.ctx.check(self.lib.hfpx_batch_solve(self.ctx.handle, self.handle, int(n_refine), 1 if use_graph else 0))

    def solution(self, i):
        """(x, backward error, info) of crack i."""
        x = np.zeros(int(self.n[i]))
        nn, berr, info = ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_int64(0)
        self.ctx.check(self.lib.hfpx_batch_solution(self.ctx.handle, self.handle, int(i), _np_ptr(x),
                                                    ctypes.byref(nn), ctypes.byref(berr), ctypes.byref(info)))
        return x, berr.value, info.value

    def rhs(self, i):
        b = np.zeros(int(self.n[i]))
        self.ctx.check(self.lib.hfpx_batch_rhs(self.ctx.handle, self.handle, int(i), _np_ptr(b)))
        return b

    def matrix(self, i):
        n = int(self.n[i])
        A = np.zeros((n, n), order="F")
        self.ctx.check(self.lib.hfpx_batch_matrix(self.ctx.handle, self.handle, int(i), _np_ptr(A), None))
        return A


# --------------------------------------------------------------------------
# cohesion / friction state (src/Development/cohesion_friction.cpp)
# --------------------------------------------------------------------------

def _fc_tables(prop):
    n_mat = len(prop.f_c_param)
    if n_mat < 1:
        raise ValueError("prop.f_c_param is empty")
    par = (_lib.FcParam * n_mat)(*[_lib.FcParam(p.cr_open, p.peak_ts, p.cr_slip, p.peak_sc, p.peak_sf, p.res_sf)
                                   for p in prop.f_c_param])
    cm = None
    if prop.c_model_id is not None:
        cm = np.asfortranarray(prop.c_model_id, dtype=np.int32)
        if cm.shape != (n_mat, 2):
            raise ValueError("prop.c_model_id must be n_surface x 2 (one row per f_c_param entry)")
    return par, cm, n_mat


def _fc_device(ctx, mat_id, dd, arrays):
    torch = ctx.torch
    dd = np.asarray(dd, dtype=np.float64)
    if dd.ndim != 2 or dd.shape[1] != 3:
        raise ValueError("dd must be n_nod x 3")
    n = dd.shape[0]
    mid = np.ascontiguousarray(mat_id, dtype=np.int32).reshape(-1)
    if mid.size != n:
        raise ValueError("mat_id must have one entry per DD row")
    with torch.cuda.stream(ctx.stream):
        d_dd = torch.from_numpy(np.ascontiguousarray(dd.T)).to(ctx.device)     # (3, n): col-major n x 3
        d_mid = torch.from_numpy(mid).to(ctx.device)
        out = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).to(ctx.device)
               for a in arrays]
    for a in out:
        if a.numel() != n:
            raise ValueError("state arrays must have one entry per DD row")
    return n, d_dd, d_mid, out


def match_f_c(prop, mat_id, dd, is_dd_local, f_state, ctx=None):
    """Friction coefficients and cohesions of every node from its DD
    (n_nod x 3: shear 0, 1, opening 2, used as given; the reference's
    rotation for global DD is an empty branch, so is_dd_local changes
    nothing), material id and reached opening / slip; writes
    f_state.friction_coef / slip_cohesion / open_cohesion.  Mirrors
    hfp3d::match_f_c (src/Development/cohesion_friction.cpp:22-127) on the
    device."""
    ctx = ctx or get_context()
    par, cm, n_mat = _fc_tables(prop)
    if cm is None:
        raise ValueError("prop.c_model_id is required")
    n = np.asarray(dd).shape[0]
    zeros = np.zeros(n)
    n, d_dd, d_mid, (mo, ms, fc, sc, oc) = _fc_device(ctx, mat_id, dd, [f_state.mr_open, f_state.mr_slip,
                                                                          zeros, zeros, zeros])
    ctx.enter()
    ctx.check(ctx.lib.hfpx_match_f_c(ctx.handle, n, _ptr(d_dd), n, _ptr(d_mid), _np_ptr(cm), par, n_mat, _ptr(mo),
                                     _ptr(ms), _ptr(fc), _ptr(sc), _ptr(oc)))
    ctx.leave()
    f_state.friction_coef = fc.cpu().numpy()
    f_state.slip_cohesion = sc.cpu().numpy()
    f_state.open_cohesion = oc.cpu().numpy()
    return f_state


def update_mrd(prop, mat_id, dd, is_dd_local, f_state, ctx=None):
    """Largest relative opening and slip reached (f_state.mr_open /
    mr_slip), updated from the current DD.  Mirrors hfp3d::update_mrd
    (src/Development/cohesion_friction.cpp:131-189) on the device."""
    ctx = ctx or get_context()
    par, _, n_mat = _fc_tables(prop)
    n, d_dd, d_mid, (mo, ms) = _fc_device(ctx, mat_id, dd, [f_state.mr_open, f_state.mr_slip])
    ctx.enter()
    ctx.check(ctx.lib.hfpx_update_mrd(ctx.handle, n, _ptr(d_dd), n, _ptr(d_mid), par, n_mat, _ptr(mo), _ptr(ms)))
    ctx.leave()
    f_state.mr_open = mo.cpu().numpy()
    f_state.mr_slip = ms.cpu().numpy()
    return f_state


# --------------------------------------------------------------------------
# cohesion / friction iteration (src/Development/c_f_iteration.cpp)
# --------------------------------------------------------------------------

class CfSystem:
    """Resident state of the volume-control cohesion / friction iteration:
    the (n_dof+1)^2 FP64 volume-control matrix (device, kept by reference),
    the device mesh of its original DoF handle, and the truncated system
    with the FP32 factors of the last active set."""

    def __init__(self, orig_matrix, device_mesh, ctx=None):
        self.ctx = ctx or get_context()
        n1, lda = _check_colmajor(orig_matrix)
        if orig_matrix.dtype != self.ctx.torch.float64 or n1 != device_mesh.n_dof + 1:
            raise ValueError("orig_matrix must be the FP64 (n_dof+1)^2 volume-control matrix of the device mesh")
        self.matrix, self.mesh = orig_matrix, device_mesh
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.hfpx_cf_create(self.ctx.handle, device_mesh.handle, _ptr(orig_matrix), lda,
                                                   ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            if _hip_calls_allowed():
                self.ctx.lib.hfpx_cf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def vc_cf_iteration(m_data_p, orig_vc_sys, orig_dof_h, prop, n_par, load, new_vol, m_data, ctx=None, tol=1e-12,
                    max_outer=10, report=None):
    """One step of the volume-control scheme with cohesion and friction on a
    pre-existing mesh; returns the residual |DD and pressure increments|_1 +
    |traction adjustments|_1 and updates m_data in place: dd, pp (the
    pressure of the collocation points that opened), dof_h_dd (the new
    active handle) and frac_state (friction / cohesion of the entry DD).
    orig_vc_sys.matrix is the resident FP64 volume-control matrix of
    orig_dof_h (make_3dbem_matrix_vc); the device mesh and the cached FP32
    factors of the truncated system are kept on orig_vc_sys between steps.
    load.inj_loc[0] is the injection node (element, local node).  m_data_p
    (the previous time step) is accepted and unused, as in the reference.
    Mirrors hfp3d::vc_cf_iteration (src/Development/c_f_iteration.cpp:27-436)
    on the device, with the traction adjustment entering the system as
    -R^T dt in all three components and the opening of a collocation point
    taken in the element frame (see hfpx_vc_cf_iteration)."""
    ctx = ctx or get_context()
    torch = ctx.torch
    cf = getattr(orig_vc_sys, "_cf", None)
    if cf is None or cf.matrix is not orig_vc_sys.matrix:
        dm = DeviceMesh(m_data.mesh, orig_dof_h, n_par.beta, ctx, dd_local=n_par.is_dd_local)
        dm.set_vertex_weights(getattr(m_data.mesh, "vert_wts", None))
        cf = CfSystem(orig_vc_sys.matrix, dm, ctx)
        orig_vc_sys._cf = cf
    cf.mesh.set_dd_local(n_par.is_dd_local)
    n_el = cf.mesh.n_el
    n_nod = 6 * n_el
    par, cm, n_mat = _fc_tables(prop)
    if cm is None:
        raise ValueError("prop.c_model_id is required")
    dd = np.asarray(m_data.dd, dtype=np.float64)
    if dd.shape != (n_nod, 3):
        raise ValueError("m_data.dd must be (6 n_el) x 3")
    fs = m_data.frac_state if m_data.frac_state is not None else Frac_State_T.zeros(n_nod)
    if load.inj_loc is None:
        raise ValueError("load.inj_loc (injection element, local node) is required")
    inj = np.asarray(load.inj_loc, dtype=np.int64).reshape(-1, 2)[0]
    dh = np.array(np.asfortranarray(m_data.dof_h_dd.dof_h, dtype=np.int64), order="F")
    if dh.shape != (n_el, 18):
        raise ValueError("m_data.dof_h_dd.dof_h must be n_el x 18")
    si = np.ascontiguousarray(load.s_inf, dtype=np.float64).reshape(6)
    with torch.cuda.stream(ctx.stream):
        d_dd = torch.from_numpy(np.ascontiguousarray(dd.T)).to(ctx.device)  # (3, n): col-major n x 3
        d_pp = torch.from_numpy(np.ascontiguousarray(m_data.pp, dtype=np.float64).reshape(-1)).to(ctx.device)
        d_mid = torch.from_numpy(np.ascontiguousarray(m_data.mat_id, dtype=np.int32).reshape(-1)).to(ctx.device)
        d_mo = torch.from_numpy(np.ascontiguousarray(fs.mr_open, dtype=np.float64).reshape(-1)).to(ctx.device)
        d_ms = torch.from_numpy(np.ascontiguousarray(fs.mr_slip, dtype=np.float64).reshape(-1)).to(ctx.device)
        fc = torch.zeros(n_nod, dtype=torch.float64, device=ctx.device)
        sc = torch.zeros_like(fc)
        oc = torch.zeros_like(fc)
    for t in (d_pp, d_mid, d_mo, d_ms):
        if t.numel() != n_nod:
            raise ValueError("pp, mat_id and the fracture state need one entry per node (6 n_el)")
    nd = ctypes.c_int64(int(m_data.dof_h_dd.n_dof))
    rep = report if report is not None else _lib.CfReport()
    ctx.enter()
    rc = ctx.lib.hfpx_vc_cf_iteration(ctx.handle, cf.handle, _ptr(d_mid), _np_ptr(cm), par, n_mat, _np_ptr(si),
                                      int(inj[0]), int(inj[1]), float(new_vol), _ptr(d_dd), n_nod, _ptr(d_pp),
                                      _ptr(d_mo), _ptr(d_ms), _ptr(fc), _ptr(sc), _ptr(oc), _np_ptr(dh),
                                      ctypes.byref(nd), float(tol), int(max_outer), ctypes.byref(rep))
    ctx.leave()
    ctx.check(rc)
    m_data.dd = d_dd.cpu().numpy().T.copy()
    m_data.pp = d_pp.cpu().numpy()
    m_data.dof_h_dd = DoF_Handle_T(dof_h=dh, n_dof=nd.value)
    fs.friction_coef, fs.slip_cohesion, fs.open_cohesion = fc.cpu().numpy(), sc.cpu().numpy(), oc.cpu().numpy()
    m_data.frac_state = fs
    return rep.residual
