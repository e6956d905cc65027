"""Python host for the time march: the ADVNCE drop-in driven through the C ABI.

``Solver`` plays the role of the reference's DG_LIN_CONVEC + ADVNCE pair
(src/dg_lin_convec.f90:96-206, src/advnce.f90:3-492): it hands the mesh,
velocity grid and deck keys to the library, advances f on the device and
reads back the reference's observables (residual log, point fields,
forces).  ``Setup`` is the same preprocessing without a GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _mesh_arrays(mesh):
    return (np.ascontiguousarray(mesh.intma, dtype=np.int32),
            np.ascontiguousarray(mesh.coord, dtype=np.float64),
            np.ascontiguousarray(mesh.bsido, dtype=np.int32))


class Setup:
    """Host-only setup (no HIP call): the reference's side table, geometry,
    normals, lumped mass, velocity grid and the library's device tables."""

    def __init__(self, case, lookup=_lib.LOOKUP_CORRECTED, precision=_lib.DGBGK_FP64,
                 order=_lib.ORDER_REFERENCE):
        L = lib()
        m = case.mesh
        self.E, self.P, self.B = m.nelem, m.npoin, m.nboun
        self._arrs = _mesh_arrays(m)
        self._eta = np.ascontiguousarray(case.eta, dtype=np.float64)
        self._w = np.ascontiguousarray(case.w, dtype=np.float64)
        self._prm = _lib.make_params(case.prm, precision, lookup, order)
        h = C.c_void_p()
        check(L.dgbgk_setup_create(self.E, self.P, self.B, *self._arrs, case.rorder,
                                   case.torder, self._eta, self._w, C.byref(self._prm),
                                   C.byref(h)))
        self.h = h
        s, v, t, w, r = (C.c_int() for _ in range(5))
        check(L.dgbgk_setup_sizes(h, C.byref(s), C.byref(v), C.byref(t), C.byref(w),
                                  C.byref(r)))
        self.S, self.V, self.ntiles, self.nwall, self.ring_separable = (
            s.value, v.value, t.value, w.value, bool(r.value))

    def close(self):
        if getattr(self, "h", None):
            lib().dgbgk_setup_destroy(self.h)
            self.h = None

    __del__ = close

    def tables(self) -> dict:
        t = _alloc_tables(self.S, self.E, self.B, self.V)
        sw, dc = C.c_double(), C.c_double()
        check(lib().dgbgk_setup_tables_host(self.h, t["iside"], t["nx"], t["ny"], t["el"],
                                            t["geome"], t["relen"], t["mmat"], t["rsido"],
                                            t["vcord"], C.byref(sw), C.byref(dc)))
        return _shape_tables(t, self.S, self.E, self.B, self.V, sw.value, dc.value)

    def reflections(self) -> dict:
        """GETBOU's reflection table: side (NW,) 1-based boundary sides; i1,
        i2 (NW, V) 1-based bracketing ordinates, 0 where not reflected; zd1,
        zd2 (NW, V) the ZETA bracket width and offset."""
        NW, V = self.nwall, self.V
        side = np.zeros(max(NW, 1), np.int32)
        i1, i2 = np.zeros(max(NW * V, 1), np.int32), np.zeros(max(NW * V, 1), np.int32)
        zd1, zd2 = np.zeros(max(NW * V, 1)), np.zeros(max(NW * V, 1))
        check(lib().dgbgk_setup_reflections(self.h, side, i1, i2, zd1, zd2))
        shp = (NW, V)
        return dict(side=side[:NW], i1=i1[:NW * V].reshape(shp), i2=i2[:NW * V].reshape(shp),
                    zd1=zd1[:NW * V].reshape(shp), zd2=zd2[:NW * V].reshape(shp))


def _alloc_tables(S, E, B, V):
    return dict(iside=np.zeros(8 * S, np.int32), nx=np.zeros(S), ny=np.zeros(S),
                el=np.zeros(S), geome=np.zeros(7 * E), relen=np.zeros(E),
                mmat=np.zeros(3 * E), rsido=np.zeros(max(3 * B, 1)), vcord=np.zeros(3 * V))


def _shape_tables(t, S, E, B, V, sw, dc):
    return dict(iside=t["iside"].reshape(S, 8), nx=t["nx"], ny=t["ny"], el=t["el"],
                geome=t["geome"].reshape(E, 7), relen=t["relen"],
                mmat=t["mmat"].reshape(E, 3), rsido=t["rsido"][:3 * B].reshape(B, 3),
                vcord=t["vcord"].reshape(V, 3), sumweight=sw, dtec=dc)


class Solver:
    """Device-resident time march for one Case.

    precision: _lib.DGBGK_FP64 or DGBGK_FP32 (FP32 f/flux/relaxation, FP64
    moments and reductions); lookup: reference or corrected wall ring rule;
    order: ORDER_REFERENCE (in-place sweep) or ORDER_JACOBI.
    """

    def __init__(self, case, precision=_lib.DGBGK_FP64, lookup=_lib.LOOKUP_CORRECTED,
                 order=_lib.ORDER_REFERENCE, vchunk: int = 0):
        L = lib()
        m = case.mesh
        self.case = case
        self.E, self.P, self.B = m.nelem, m.npoin, m.nboun
        self.V = case.nvel
        self._arrs = _mesh_arrays(m)
        self._eta = np.ascontiguousarray(case.eta, dtype=np.float64)
        self._w = np.ascontiguousarray(case.w, dtype=np.float64)
        self._prm = _lib.make_params(case.prm, precision, lookup, order, vchunk)
        h = C.c_void_p()
        check(L.dgbgk_create(self.E, self.P, self.B, *self._arrs, case.rorder, case.torder,
                             self._eta, self._w, C.byref(self._prm), C.byref(h)))
        self.h = h
        s = C.c_int()
        check(L.dgbgk_sizes(h, C.byref(s), None, None, None, None, None))
        self.S = s.value
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            lib().dgbgk_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def itime(self) -> int:
        it = C.c_int()
        check(lib().dgbgk_sizes(self.h, None, None, None, None, None, C.byref(it)))
        return it.value

    def step(self, n: int = 1, records: bool = True):
        """Advance n steps; returns per-step records (dicts) when asked."""
        if not records:
            check(lib().dgbgk_step(self.h, n, None))
            return None
        info = (_lib.StepInfo * n)()
        check(lib().dgbgk_step(self.h, n, C.cast(info, C.c_void_p)))
        return [dict(sumres=i.sumres, logres=i.logres, dte=i.dte, mass=i.mass, lift=i.lift,
                     drag=i.drag, change=i.change) for i in info]

    def get_f(self) -> np.ndarray:
        """f in the reference layout DISNF(3,V,E) as an (E,V,3) array."""
        f = np.empty(3 * self.E * self.V)
        check(lib().dgbgk_get_f(self.h, f))
        return f.reshape(self.E, self.V, 3)

    def set_f(self, f: np.ndarray):
        check(lib().dgbgk_set_f(self.h, np.ascontiguousarray(f, dtype=np.float64).ravel()))

    def seed(self, rho, u, v, temp):
        """Per-point Maxwellian seed (the warm start's entry point)."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, u, v, temp)]
        check(lib().dgbgk_seed_maxwellian(self.h, *a))

    def node_fields(self) -> dict:
        a = [np.empty(3 * self.E) for _ in range(5)]
        check(lib().dgbgk_node_fields(self.h, *a))
        return dict(zip(("disnd", "disux", "disuy", "disps", "mu"),
                        (x.reshape(self.E, 3) for x in a)))

    def point_fields(self) -> dict:
        a = [np.empty(self.P) for _ in range(6)]
        check(lib().dgbgk_point_fields(self.h, *a))
        return dict(zip(("nd", "rho", "uvel", "vvel", "ps", "temp"), a))

    def forces(self):
        lft, drg = C.c_double(), C.c_double()
        check(lib().dgbgk_forces(self.h, C.byref(lft), C.byref(drg)))
        return lft.value, drg.value

    def wall_eta(self) -> np.ndarray:
        e = np.empty(max(self.B, 1))
        check(lib().dgbgk_wall_eta(self.h, e))
        return e[: self.B]

    def tables(self) -> dict:
        t = _alloc_tables(self.S, self.E, self.B, self.V)
        sw, dc = C.c_double(), C.c_double()
        check(lib().dgbgk_setup_tables(self.h, t["iside"], t["nx"], t["ny"], t["el"],
                                       t["geome"], t["relen"], t["mmat"], t["rsido"],
                                       t["vcord"], C.byref(sw), C.byref(dc)))
        return _shape_tables(t, self.S, self.E, self.B, self.V, sw.value, dc.value)

    def timing(self, enable: bool = True):
        check(lib().dgbgk_timing(self.h, 1 if enable else 0))

    def kernel_time(self, kernel: int):
        ms, cnt = C.c_double(), C.c_long()
        check(lib().dgbgk_kernel_time(self.h, kernel, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def sync(self):
        check(lib().dgbgk_sync(self.h))

    def f_stats(self) -> dict:
        """min and max of the finite values of f and the count of non-finite
        ones, computed on the device over every node and ordinate."""
        lo, hi, bad = C.c_double(), C.c_double(), C.c_longlong()
        check(lib().dgbgk_f_stats(self.h, C.byref(lo), C.byref(hi), C.byref(bad)))
        return {"min": lo.value, "max": hi.value, "nonfinite": bad.value}

    def f_stats_noalloc(self) -> dict:
        """f_stats through the allocation-free entry point (fixed per-GPU
        scratch; at most 4096 ordinates)."""
        lo, hi, bad = C.c_double(), C.c_double(), C.c_longlong()
        check(lib().dgbgk_f_stats_noalloc(self.h, C.byref(lo), C.byref(hi), C.byref(bad)))
        return {"min": lo.value, "max": hi.value, "nonfinite": bad.value}

    def bytes_per_value(self) -> int:
        b = C.c_int()
        check(lib().dgbgk_dof_bytes(self.h, C.byref(b)))
        return b.value

    def launch_config(self) -> dict:
        v = [C.c_int() for _ in range(6)]
        check(lib().dgbgk_launch_config(self.h, *(C.byref(x) for x in v)))
        return dict(zip(("slots", "depth", "vchunk", "nvchunks", "ntiles", "mom_split"),
                        (x.value for x in v)))
