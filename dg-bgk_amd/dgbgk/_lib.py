"""ctypes declarations of the C ABI in include/dgbgk.h.

Loading fails loudly when libdgbgk.so has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# DGBGK_LIB: another build of the same library (for comparing builds)
LIB_PATH = os.environ.get("DGBGK_LIB") or os.path.join(PKG, "libdgbgk.so")

DGBGK_FP64, DGBGK_FP32 = 0, 1
LOOKUP_REFERENCE, LOOKUP_CORRECTED = 0, 1
ORDER_REFERENCE, ORDER_JACOBI = 0, 1

ERRORS = {
    -1: "DGBGK_ERR_ARG", -2: "DGBGK_ERR_MXSID", -3: "DGBGK_ERR_BOUNDARY_SEARCH",
    -4: "DGBGK_ERR_WALL_SIDE", -5: "DGBGK_ERR_SUMBOTTOM", -6: "DGBGK_ERR_REFLECT_LOOKUP",
    -7: "DGBGK_ERR_MESH", -8: "DGBGK_ERR_UNSUPPORTED", -9: "DGBGK_ERR_HIP",
    -10: "DGBGK_ERR_NOMEM", -11: "DGBGK_ERR_TILE", -12: "DGBGK_ERR_STATE",
}

# every entry point declared in include/dgbgk.h
EXPORTS = [
    "dgbgk_create", "dgbgk_destroy", "dgbgk_last_error", "dgbgk_step", "dgbgk_get_f",
    "dgbgk_set_f", "dgbgk_seed_maxwellian", "dgbgk_node_fields", "dgbgk_point_fields",
    "dgbgk_forces", "dgbgk_wall_eta", "dgbgk_sizes", "dgbgk_setup_tables",
    "dgbgk_setup_create", "dgbgk_setup_sizes", "dgbgk_setup_tables_host",
    "dgbgk_setup_reflections",
    "dgbgk_setup_destroy", "dgbgk_timing", "dgbgk_kernel_time", "dgbgk_sync",
    "dgbgk_dof_bytes", "dgbgk_launch_config", "dgbgk_f_stats", "dgbgk_f_stats_noalloc",
    "dgbgk_mlp_create",
    "dgbgk_mlp_destroy", "dgbgk_mlp_last_error",
    "dgbgk_mlp_nparams", "dgbgk_mlp_set_params", "dgbgk_mlp_get_params", "dgbgk_mlp_forward",
    "dgbgk_mlp_loss_grad", "dgbgk_mlp_set_data", "dgbgk_mlp_train",
]


class DgbgkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("rv", C.c_double), ("csafm", C.c_double), ("alpha", C.c_double),
                ("rgas", C.c_double), ("d", C.c_double), ("m", C.c_double),
                ("t1", C.c_double), ("p1", C.c_double), ("u0", C.c_double),
                ("v0", C.c_double), ("cinf", C.c_double * 4), ("inf", C.c_int),
                ("immat", C.c_int), ("precision", C.c_int), ("lookup", C.c_int),
                ("order", C.c_int), ("vchunk", C.c_int)]


class StepInfo(C.Structure):
    _fields_ = [("sumres", C.c_double), ("logres", C.c_double), ("dte", C.c_double),
                ("mass", C.c_double), ("lift", C.c_double), ("drag", C.c_double),
                ("change", C.c_double)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_L = None


def _opt(t):
    """ndpointer that also accepts None."""
    class _Opt(t):  # type: ignore[misc, valid-type]
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return t.from_param(obj)
    return _Opt


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python __graft_entry__.py build` "
            "(there is no CPU fallback for the time march)")
    L = C.CDLL(LIB_PATH)
    odp, oip = _opt(_dp), _opt(_ip)
    vp = C.c_void_p
    L.dgbgk_last_error.restype = C.c_char_p
    L.dgbgk_create.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _ip, C.c_int, C.c_int,
                               _dp, _dp, C.POINTER(Params), C.POINTER(vp)]
    L.dgbgk_destroy.argtypes = [vp]
    L.dgbgk_destroy.restype = None
    L.dgbgk_step.argtypes = [vp, C.c_int, C.c_void_p]
    L.dgbgk_get_f.argtypes = [vp, _dp]
    L.dgbgk_set_f.argtypes = [vp, _dp]
    L.dgbgk_seed_maxwellian.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.dgbgk_node_fields.argtypes = [vp] + [odp] * 5
    L.dgbgk_point_fields.argtypes = [vp] + [odp] * 6
    L.dgbgk_forces.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.dgbgk_wall_eta.argtypes = [vp, _dp]
    L.dgbgk_sizes.argtypes = [vp] + [C.POINTER(C.c_int)] * 6
    L.dgbgk_setup_tables.argtypes = [vp, oip] + [odp] * 8 + [C.POINTER(C.c_double)] * 2
    L.dgbgk_setup_create.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _ip, C.c_int,
                                     C.c_int, _dp, _dp, C.POINTER(Params), C.POINTER(vp)]
    L.dgbgk_setup_sizes.argtypes = [vp] + [C.POINTER(C.c_int)] * 5
    L.dgbgk_setup_tables_host.argtypes = [vp, oip] + [odp] * 8 + [C.POINTER(C.c_double)] * 2
    L.dgbgk_setup_reflections.argtypes = [vp, oip, oip, oip, odp, odp]
    L.dgbgk_setup_destroy.argtypes = [vp]
    L.dgbgk_setup_destroy.restype = None
    L.dgbgk_timing.argtypes = [vp, C.c_int]
    L.dgbgk_kernel_time.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long)]
    L.dgbgk_sync.argtypes = [vp]
    L.dgbgk_dof_bytes.argtypes = [vp, C.POINTER(C.c_int)]
    L.dgbgk_f_stats.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_longlong)]
    L.dgbgk_f_stats_noalloc.argtypes = L.dgbgk_f_stats.argtypes
    L.dgbgk_launch_config.argtypes = [vp] + [C.POINTER(C.c_int)] * 6
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    ofp = _opt(fp)
    L.dgbgk_mlp_create.argtypes = [C.c_int] * 5 + [C.POINTER(vp)]
    L.dgbgk_mlp_destroy.argtypes = [vp]
    L.dgbgk_mlp_destroy.restype = None
    L.dgbgk_mlp_last_error.restype = C.c_char_p
    L.dgbgk_mlp_nparams.argtypes = [vp, C.POINTER(C.c_long)]
    L.dgbgk_mlp_set_params.argtypes = [vp, fp]
    L.dgbgk_mlp_get_params.argtypes = [vp, fp]
    L.dgbgk_mlp_forward.argtypes = [vp, C.c_long, fp, fp]
    L.dgbgk_mlp_loss_grad.argtypes = [vp, C.c_int, fp, fp, C.POINTER(C.c_float), ofp]
    L.dgbgk_mlp_set_data.argtypes = [vp, C.c_long, fp, fp]
    L.dgbgk_mlp_train.argtypes = [vp, C.c_int, C.c_float, C.c_uint, ofp]
    _L = L
    return L


def check(code: int):
    if code != 0:
        raise DgbgkError(code, lib().dgbgk_last_error().decode())


def make_params(prm: dict, precision: int = DGBGK_FP64, lookup: int = LOOKUP_CORRECTED,
                order: int = ORDER_REFERENCE, vchunk: int = 0) -> Params:
    p = Params()
    for k in ("rv", "csafm", "alpha", "rgas", "d", "m", "t1", "p1", "u0", "v0"):
        setattr(p, k, float(prm[k]))
    for i in range(4):
        p.cinf[i] = float(prm["cinf"][i])
    p.inf = int(prm.get("inf", 1))
    p.immat = int(prm.get("immat", 1))
    p.precision, p.lookup, p.order, p.vchunk = int(precision), int(lookup), int(order), int(vchunk)
    return p
