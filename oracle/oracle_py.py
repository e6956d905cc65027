"""ctypes access to the FP64 C oracle and to the reference's own routines.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OraParams(C.Structure):
    _fields_ = [("rv", C.c_double), ("csafm", C.c_double), ("alpha", C.c_double),
                ("rgas", C.c_double), ("d", C.c_double), ("m", C.c_double),
                ("t1", C.c_double), ("p1", C.c_double), ("u0", C.c_double),
                ("v0", C.c_double), ("cinf", C.c_double * 4), ("inf", C.c_int),
                ("lookup", C.c_int), ("jacobi", C.c_int), ("faithful_scans", C.c_int),
                ("mutant", C.c_int), ("immat", C.c_int)]


class OraStepInfo(C.Structure):
    _fields_ = [("sumres", C.c_double), ("logres", C.c_double), ("dte", C.c_double),
                ("mass", C.c_double), ("lift", C.c_double), ("drag", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.ora_create.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _ip, C.c_int,
                                 C.c_int, _dp, _dp, C.POINTER(OraParams),
                                 C.POINTER(C.c_void_p)]
        L.ora_create.restype = C.c_int
        L.ora_destroy.argtypes = [C.c_void_p]
        L.ora_nside.argtypes = [C.c_void_p]
        L.ora_nvel.argtypes = [C.c_void_p]
        L.ora_iside.argtypes = [C.c_void_p]
        L.ora_iside.restype = C.POINTER(C.c_int)
        L.ora_vcord.argtypes = [C.c_void_p]
        L.ora_vcord.restype = C.POINTER(C.c_double)
        L.ora_sumweight.argtypes = [C.c_void_p]
        L.ora_sumweight.restype = C.c_double
        L.ora_get_f.argtypes = [C.c_void_p, _dp]
        L.ora_cmmat.argtypes = [C.c_void_p, _dp]
        L.ora_cmmat.restype = None
        L.ora_conmat.argtypes = [C.c_int, _dp, _dp]
        L.ora_conmat.restype = None
        L.ora_set_f.argtypes = [C.c_void_p, _dp]
        L.ora_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(OraStepInfo)]
        L.ora_step.restype = C.c_int
        L.ora_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.ora_set_threads.restype = C.c_int
        L.ora_set_emulation.argtypes = [C.c_void_p, C.c_int]
        L.ora_set_emulation.restype = None
        L.ora_get_node_fields.argtypes = [C.c_void_p] + [_dp] * 5
        L.ora_get_point_fields.argtypes = [C.c_void_p] + [_dp] * 6
        L.ora_get_eta.argtypes = [C.c_void_p, _dp]
        L.ora_getsid.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int]
        L.ora_getsid.restype = C.c_int
        L.ora_normal.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp]
        L.ora_getgeo.argtypes = [C.c_int, _ip, _dp, _dp]
        L.ora_getlen.argtypes = [C.c_int, _dp, _dp]
        L.ora_getmat.argtypes = [C.c_int, _dp, _dp]
        L.ora_getnor.argtypes = [C.c_int, _ip, _dp, _dp]
        L.ora_vspace.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_double)]
        L.ora_gteqnf.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                 C.c_double, _dp]
        L.ora_getfor.argtypes = [C.c_int, _ip, _dp, _dp, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.ora_inflow_value.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.ora_inflow_value.restype = C.c_double
        L.ora_prepare_step.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ora_debug_ordinate.argtypes = [C.c_void_p, C.c_int, C.c_double] + [_dp] * 5 + [
            C.POINTER(C.c_double)]
        L.ora_velocity.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        L.ora_tables.argtypes = [C.c_void_p] + [_dp] * 7 + [C.POINTER(C.c_double)]
        L.ora_cinf.argtypes = [C.c_void_p, _dp]
        _lib = L
    return _lib


MUTANT_DOWNWIND, MUTANT_SWAP_NODES, MUTANT_GETMAC_EVERY, MUTANT_RES_AFTER = 1, 2, 3, 4
MUTANT_CMMINC_RHS_JN = 5
EMU_STORE_RN, EMU_STORE_SR, EMU_PARAMS, EMU_ARITH, EMU_DTE = 1, 2, 4, 8, 16
EMU_PRED, EMU_EDGE, EMU_INC, EMU_FORM = 32, 64, 128, 256


def make_params(prm: dict, lookup: int = 1, jacobi: int = 0, faithful: int = 0,
                mutant: int = 0) -> OraParams:
    p = OraParams()
    for k in ("rv", "csafm", "alpha", "rgas", "d", "m", "t1", "p1", "u0", "v0"):
        setattr(p, k, float(prm[k]))
    for i in range(4):
        p.cinf[i] = float(prm["cinf"][i])
    p.inf = int(prm.get("inf", 1))
    p.lookup, p.jacobi, p.faithful_scans = int(lookup), int(jacobi), int(faithful)
    p.mutant = int(mutant)
    p.immat = int(prm.get("immat", 1))
    return p


class OracleRun:
    """One oracle context on a Case (see dgbgk.cases)."""

    def __init__(self, case, lookup: int = 1, jacobi: int = 0, faithful: int = 0,
                 mutant: int = 0):
        L = lib()
        m = case.mesh
        self.case = case
        self.E, self.P, self.B = m.nelem, m.npoin, m.nboun
        self.V = case.nvel
        self._intma = np.ascontiguousarray(m.intma, dtype=np.int32)
        self._coord = np.ascontiguousarray(m.coord, dtype=np.float64)
        self._bsido = np.ascontiguousarray(m.bsido, dtype=np.int32)
        self._prm = make_params(case.prm, lookup, jacobi, faithful, mutant)
        h = C.c_void_p()
        st = L.ora_create(self.E, self.P, self.B, self._intma, self._coord, self._bsido,
                          case.rorder, case.torder, np.ascontiguousarray(case.eta, float),
                          np.ascontiguousarray(case.w, float), C.byref(self._prm),
                          C.byref(h))
        if st != 0:
            raise RuntimeError(f"ora_create failed with status {st}")
        self.h = h
        self.S = L.ora_nside(h)

    def close(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    __del__ = close

    def step(self, n: int = 1):
        info = (OraStepInfo * n)()
        st = lib().ora_step(self.h, n, info)
        if st != 0:
            raise RuntimeError(f"ora_step failed with status {st}")
        return [dict(sumres=i.sumres, logres=i.logres, dte=i.dte, mass=i.mass,
                     lift=i.lift, drag=i.drag) for i in info]

    def set_threads(self, n: int):
        """Threads for step() (the same values bit for bit as one thread)."""
        if lib().ora_set_threads(self.h, int(n)) != 0:
            raise MemoryError("ora_set_threads: table allocation failed")

    def set_emulation(self, flags: int):
        """FP32 emulation of stages of the step (EMU_* flags; 0 = FP64)."""
        lib().ora_set_emulation(self.h, int(flags))

    def cmmat(self) -> np.ndarray:
        """CONMAT's matrices, shape (E, 3, 3) as [IE][JN][IN] (zeros unless
        the context was created with immat = 0)."""
        m = np.empty(9 * self.E)
        lib().ora_cmmat(self.h, m)
        return m.reshape(self.E, 3, 3)

    def get_f(self) -> np.ndarray:
        """f in the reference layout, shape (E, V, 3)."""
        f = np.empty(self.E * self.V * 3)
        lib().ora_get_f(self.h, f)
        return f.reshape(self.E, self.V, 3)

    def set_f(self, f: np.ndarray):
        lib().ora_set_f(self.h, np.ascontiguousarray(f, dtype=np.float64).ravel())

    def iside(self) -> np.ndarray:
        p = lib().ora_iside(self.h)
        return np.ctypeslib.as_array(p, shape=(self.S * 8,)).reshape(self.S, 8).copy()

    def vcord(self) -> np.ndarray:
        p = lib().ora_vcord(self.h)
        return np.ctypeslib.as_array(p, shape=(self.V * 3,)).reshape(self.V, 3).copy()

    def sumweight(self) -> float:
        return lib().ora_sumweight(self.h)

    def node_fields(self):
        a = [np.empty(3 * self.E) for _ in range(5)]
        lib().ora_get_node_fields(self.h, *a)
        return dict(zip(("disnd", "disux", "disuy", "disps", "mu"),
                        (x.reshape(self.E, 3) for x in a)))

    def point_fields(self):
        a = [np.empty(self.P) for _ in range(6)]
        lib().ora_get_point_fields(self.h, *a)
        return dict(zip(("nd", "rho", "uvel", "vvel", "ps", "temp"), a))

    def prepare_step(self) -> float:
        dte = C.c_double()
        st = lib().ora_prepare_step(self.h, C.byref(dte))
        if st != 0:
            raise RuntimeError(f"ora_prepare_step failed with status {st}")
        return dte.value

    def debug_ordinate(self, iv: int, dte: float) -> dict:
        E = self.E
        out = dict(nfo=np.empty(3 * E), umean=np.empty(E), rhs_elc=np.empty(3 * E),
                   rhs_edge=np.empty(3 * E), delun=np.empty(3 * E))
        res = C.c_double()
        st = lib().ora_debug_ordinate(self.h, iv, dte, out["nfo"], out["umean"],
                                      out["rhs_elc"], out["rhs_edge"], out["delun"],
                                      C.byref(res))
        if st != 0:
            raise RuntimeError(f"ora_debug_ordinate failed with status {st}")
        out["residual"] = res.value
        return out

    def velocity(self, iv: int):
        ux, uy = C.c_double(), C.c_double()
        lib().ora_velocity(self.h, iv, C.byref(ux), C.byref(uy))
        return ux.value, uy.value

    def tables(self) -> dict:
        S, E, B = self.S, self.E, self.B
        t = dict(nx=np.empty(S), ny=np.empty(S), el=np.empty(S), geome=np.empty(7 * E),
                 relen=np.empty(E), mmat=np.empty(3 * E), rsido=np.empty(max(3 * B, 1)))
        dtec = C.c_double()
        lib().ora_tables(self.h, t["nx"], t["ny"], t["el"], t["geome"], t["relen"],
                         t["mmat"], t["rsido"], C.byref(dtec))
        t["geome"] = t["geome"].reshape(E, 7)
        t["mmat"] = t["mmat"].reshape(E, 3)
        t["rsido"] = t["rsido"][:3 * B].reshape(B, 3)
        t["dtec"] = dtec.value
        return t

    def cinf(self):
        c = np.empty(4)
        lib().ora_cinf(self.h, c)
        return c

    def eta(self):
        e = np.empty(self.B)
        lib().ora_get_eta(self.h, e)
        return e


class RefLib:
    """The reference's own MPI-free routines (oracle/_ref/libref{64,32}.so)."""

    def __init__(self, prec: int = 64):
        path = os.path.join(HERE, "_ref", f"libref{prec}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = C.CDLL(path)
        self.dt = np.float64 if prec == 64 else np.float32
        self.ct = C.c_double if prec == 64 else C.c_float

    def arr(self, x):
        return np.ascontiguousarray(x, dtype=self.dt)

    def ptr(self, a):
        return a.ctypes.data_as(C.c_void_p)

    def call(self, name, *args):
        fn = getattr(self.L, name)
        conv = []
        for a in args:
            if isinstance(a, np.ndarray):
                conv.append(a.ctypes.data_as(C.c_void_p))
            elif isinstance(a, float):
                conv.append(self.ct(a))
            elif isinstance(a, int):
                conv.append(C.c_int(a))
            else:
                conv.append(a)
        fn.restype = None
        fn(*conv)
