"""ctypes binding of the C ABI declared in include/hfpx3d.h.

The library is the in-tree ``hfpx3d_amd/libhfpx3d.so`` built by
``hfpx3d_amd/build.py``.  There is no fallback: if the library is missing or
cannot be loaded, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhfpx3d.so")

HFPX_OK = 0
HFPX_ERR_INVALID = 1
HFPX_ERR_NOMEM = 2
HFPX_ERR_SINGULAR = 3
HFPX_ERR_NOT_CONVERGED = 4
HFPX_ERR_HIP = 5
HFPX_F64 = 0
HFPX_F32 = 1
KNOB_GEMM_F32 = 0
KNOB_LU_NB = 1
KNOB_BGEMM_F32 = 2
KNOB_PANEL = 3
KNOB_X3_BAND = 4
KNOB_IR_FALLBACK = 5
KNOB_TEST_PANEL_TIMEOUT = 6
KNOB_GEMM_F64 = 7
KNOB_LU_LOOKAHEAD = 8
KNOB_IR_MODE = 9
GEMM_F32_VARIANTS = 18

KERNEL_CLASSES = {"assemble": 0, "gemm_f64": 1, "gemm_f32": 2, "panel": 3, "laswp": 4, "trsm": 5, "stress": 6,
                  "vc": 7}

_c_void_p = ctypes.c_void_p
_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_double = ctypes.c_double


class FcParam(ctypes.Structure):
    """hfpx_fc_param: friction / cohesion parameters of one surface material."""
    _fields_ = [("cr_open", ctypes.c_double), ("peak_ts", ctypes.c_double), ("cr_slip", ctypes.c_double),
                ("peak_sc", ctypes.c_double), ("peak_sf", ctypes.c_double), ("res_sf", ctypes.c_double)]
_pp = ctypes.POINTER(ctypes.c_void_p)


class IRReport(ctypes.Structure):
    """Mirror of struct hfpx_ir_report."""
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("inner_iterations", ctypes.c_int),
        ("backward_error", ctypes.c_double),
        ("converged", ctypes.c_int),
        ("fell_back", ctypes.c_int),
    ]


class CfReport(ctypes.Structure):
    """Mirror of struct hfpx_cf_report."""
    _fields_ = [
        ("n_dof", ctypes.c_int64),
        ("refactored", ctypes.c_int),
        ("pressure", ctypes.c_double),
        ("residual", ctypes.c_double),
        ("ir", IRReport),
    ]


# name -> (restype, argtypes); must match include/hfpx3d.h exactly
SIGNATURES = {
    "hfpx_create": (_c_int, [_pp, _c_int, _c_void_p]),
    "hfpx_destroy": (_c_int, [_c_void_p]),
    "hfpx_status_string": (ctypes.c_char_p, [_c_int]),
    "hfpx_last_error": (ctypes.c_char_p, [_c_void_p]),
    "hfpx_stream": (_c_void_p, [_c_void_p]),
    "hfpx_synchronize": (_c_int, [_c_void_p]),
    "hfpx_mesh_create": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_i64, _c_void_p, _c_i64, _c_i64,
                                  _c_void_p, _c_i64, _c_double, _pp]),
    "hfpx_mesh_destroy": (_c_int, [_c_void_p]),
    "hfpx_mesh_n_dof": (_c_i64, [_c_void_p]),
    "hfpx_mesh_set_dd_local": (_c_int, [_c_void_p, _c_int]),
    "hfpx_mesh_create_npy": (_c_int, [_c_void_p, ctypes.c_char_p, ctypes.c_char_p, _c_i64, _c_int, _c_double,
                                      _c_void_p]),
    "hfpx_mesh_set_vertex_weights": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_i64]),
    "hfpx_match_f_c": (_c_int, [_c_void_p, _c_i64, _c_void_p, _c_i64, _c_void_p, _c_void_p, _c_void_p, _c_i64,
                                _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "hfpx_update_mrd": (_c_int, [_c_void_p, _c_i64, _c_void_p, _c_i64, _c_void_p, _c_void_p, _c_i64, _c_void_p,
                                 _c_void_p]),
    "hfpx_assemble": (_c_int, [_c_void_p, _c_void_p, _c_double, _c_double, _c_void_p, _c_i64]),
    "hfpx_assemble_mixed": (_c_int, [_c_void_p, _c_void_p, _c_double, _c_double, _c_void_p, _c_i64,
                                     _c_void_p, _c_i64]),
    "hfpx_rhs_sinf": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_i64]),
    "hfpx_assemble_vc": (_c_int, [_c_void_p, _c_void_p, _c_double, _c_double, _c_void_p, _c_i64, _c_void_p,
                                  _c_i64]),
    "hfpx_vc_truncate": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_i64, _c_void_p, _c_void_p, _c_i64, _c_i64,
                                  _c_i64, _c_void_p, _c_i64, _c_double, _c_void_p, _c_i64, _c_void_p, _c_i64,
                                  _c_void_p]),
    "hfpx_stress": (_c_int, [_c_void_p, _c_void_p, _c_double, _c_double, _c_void_p, _c_i64, _c_void_p, _c_i64,
                             _c_void_p, _c_void_p, _c_i64]),
    "hfpx_getrf": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_i64, _c_i64, _c_void_p, ctypes.POINTER(_c_i64)]),
    "hfpx_getrs": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_i64, _c_void_p, _c_i64, _c_void_p, _c_i64, _c_i64]),
    "hfpx_solve_ir": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_i64, _c_void_p, _c_i64, _c_void_p,
                               _c_i64, _c_void_p, _c_i64, _c_i64, _c_double, _c_int, _c_int,
                               ctypes.POINTER(IRReport)]),
    "hfpx_profile": (_c_int, [_c_void_p, _c_int]),
    "hfpx_profile_classes": (_c_int, [_c_void_p, ctypes.c_uint32]),
    "hfpx_profile_read": (_c_int, [_c_void_p, _c_int, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_double),
                                   ctypes.POINTER(_c_double), ctypes.POINTER(_c_double)]),
    "hfpx_gemm_sub": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_i64, _c_void_p, _c_i64, _c_void_p, _c_i64,
                               _c_i64, _c_i64, _c_i64]),
    "hfpx_tune_set": (_c_int, [_c_void_p, _c_int, _c_int]),
    "hfpx_batch_create": (_c_int, [_c_void_p, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_double,
                                   _c_void_p, _c_void_p, _c_void_p, _c_i64, ctypes.POINTER(_c_void_p)]),
    "hfpx_batch_destroy": (_c_int, [_c_void_p]),
    "hfpx_batch_assemble": (_c_int, [_c_void_p, _c_void_p]),
    "hfpx_batch_solve": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int]),
    "hfpx_batch_solution": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_void_p, ctypes.POINTER(_c_i64),
                                     ctypes.POINTER(_c_double), ctypes.POINTER(_c_i64)]),
    "hfpx_batch_rhs": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_void_p]),
    "hfpx_batch_matrix": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_void_p, ctypes.POINTER(_c_i64)]),
    "hfpx_tune_get": (_c_int, [_c_void_p, _c_int, ctypes.POINTER(_c_int)]),
    "hfpx_local_submatrix": (_c_int, [_c_void_p, _c_int, _c_double, _c_double, _c_i64, _c_void_p, _c_void_p,
                                      _c_void_p, _c_void_p, _c_void_p]),
    "hfpx_stress_noalloc": (_c_int, [_c_void_p, _c_void_p, _c_double, _c_double, _c_void_p, _c_i64, _c_void_p,
                                     _c_i64, _c_void_p, _c_void_p, _c_i64]),
    "hfpx_cf_create": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_i64, ctypes.POINTER(_c_void_p)]),
    "hfpx_cf_destroy": (_c_int, [_c_void_p]),
    "hfpx_vc_cf_iteration": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_i64, _c_void_p,
                                      _c_i64, _c_i64, _c_double, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                      _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                      ctypes.POINTER(_c_i64), _c_double, _c_int, ctypes.POINTER(CfReport)]),
}

_lib = None


def load():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"HIP library {LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HfpxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"hfpx status {status}: {msg}")
        self.status = status


def check(rc, ctx=None):
    if rc != HFPX_OK:
        lib = load()
        msg = lib.hfpx_status_string(rc).decode()
        if ctx is not None:
            detail = lib.hfpx_last_error(ctx).decode()
            if detail:
                msg = f"{msg}: {detail}"
        raise HfpxError(rc, msg)
