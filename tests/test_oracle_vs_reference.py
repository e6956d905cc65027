"""The C oracle against the reference's own routines (oracle/_ref, compiled
from the reference's sources with -fdefault-real-8), bit for bit.

Covers GETSID, NORMAL, GETGEO, GETLEN, GETMAT, CONMAT, GETNOR, VSPACE,
ALOTIM, ALOTIM2, GTEQNF, GETFLA+GETELC, GETBOU (all branches, both ring
rules and the in-place reads of already-advanced ordinates), GETINC+ADTHEM,
CMMINC+ADTHEM, INFLOW and GETFOR.  The routines that need MPI (FLUCON2, COMPMU, MACROS, the
interior-side loop of EDGFLXOPT) are pinned in test_oracle_closed_forms.py.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import have_ref
from small_cases import GOLDEN, c1, cylinder_lobatto10, cylinder_small, lobatto10_walls

pytestmark = pytest.mark.skipif(not have_ref(64), reason="oracle/_ref not built")


def _ref(prec=64):
    from oracle_py import RefLib
    return RefLib(prec)


def _ip(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _ref_setup(R, mesh):
    E, P, B = mesh.nelem, mesh.npoin, mesh.nboun
    intma, coord, bsido = _ip(mesh.intma), np.ascontiguousarray(mesh.coord), _ip(mesh.bsido)
    mx = 2 * (E + B)
    iside = np.zeros(8 * mx, np.int32)
    ns = np.zeros(1, np.int32)
    R.L.ref_getsid(C.c_int(E), C.c_int(P), _ptr(intma), _ptr(iside), C.c_int(mx), _ptr(ns))
    S = int(ns[0])
    iside = np.ascontiguousarray(iside[:8 * S])
    nx, ny, el = np.zeros(S), np.zeros(S), np.zeros(S)
    R.L.ref_normal(C.c_int(S), C.c_int(P), _ptr(iside), _ptr(coord), _ptr(nx), _ptr(ny), _ptr(el))
    geome = np.zeros(7 * E)
    R.L.ref_getgeo(C.c_int(E), C.c_int(P), _ptr(intma), _ptr(coord), _ptr(geome))
    relen, mmat = np.zeros(E), np.zeros(3 * E)
    R.L.ref_getlen(C.c_int(E), _ptr(geome), _ptr(relen))
    R.L.ref_getmat(C.c_int(E), _ptr(geome), _ptr(mmat))
    rsido = np.zeros(3 * max(B, 1))
    R.L.ref_getnor(C.c_int(B), C.c_int(P), _ptr(bsido), _ptr(coord), _ptr(rsido))
    return dict(iside=iside.reshape(S, 8), nx=nx, ny=ny, el=el, geome=geome.reshape(E, 7),
                relen=relen, mmat=mmat.reshape(E, 3), rsido=rsido[:3 * B].reshape(B, 3))


@pytest.mark.parametrize("make", [c1, cylinder_small, lobatto10_walls])
def test_mesh_preprocessing_bit_exact(make):
    from oracle_py import OracleRun
    case = make()
    R = _ref()
    ref = _ref_setup(R, case.mesh)
    o = OracleRun(case)
    t = o.tables()
    assert np.array_equal(o.iside(), ref["iside"])
    for k in ("nx", "ny", "el", "geome", "relen", "mmat", "rsido"):
        assert np.array_equal(t[k], ref[k]), k
    # ALOTIM (src/alotim.f90) with MXSPD = rv
    dte = C.c_double()
    R.L.ref_alotim(C.c_int(case.mesh.nelem), C.c_double(case.prm["csafm"]), _ptr(ref["relen"]),
                   C.c_double(case.prm["rv"]), C.byref(dte))
    assert dte.value == t["dtec"]


def test_side_table_invariants():
    """GETSID invariants: S = (3E + B)/2; interior sides have p1 < p2 and two
    elements; boundary sides have ISIDE(4) = 0 and match BSIDO's sides."""
    from oracle_py import OracleRun
    for case in (c1(), cylinder_small()):
        o = OracleRun(case)
        isd = o.iside()
        E, B = case.mesh.nelem, case.mesh.nboun
        assert len(isd) == (3 * E + B) // 2
        inter = isd[:, 3] != 0
        assert np.all(isd[inter, 0] < isd[inter, 1])
        assert (~inter).sum() == B
        bs = {(int(a), int(b)) for a, b in case.mesh.bsido[:, :2]}
        assert {(int(a), int(b)) for a, b in isd[~inter, :2]} == bs


def test_vspace_lobatto40_sumweight():
    """VSPACE on the reference's Lobatto40 file with TORDER=40: the FP32
    build prints SUMWEIGHT = 79.99974 (tc/tc1/output:24); the oracle matches
    the FP64 build bit for bit."""
    from oracle_py import lib
    from dgbgk import io
    path = os.path.join(GOLDEN, "Lobatto40.txt").encode()
    for prec in (32, 64):
        R = _ref(prec)
        vc = np.zeros(3 * 40 * 40, dtype=R.dt)
        sw = R.ct()
        rr = C.c_int()
        R.L.ref_vspace(C.c_char_p(path), C.c_int(len(path)), C.c_int(40), C.c_int(40),
                       C.byref(rr), _ptr(vc), C.byref(sw))
        assert rr.value == 40
        if prec == 32:
            assert f"{sw.value:.5f}" == "79.99974"
        else:
            eta, w = io.read_lobatto(path.decode())
            vo = np.zeros(3 * 1600)
            swo = C.c_double()
            lib().ora_vspace(40, 40, eta, w, vo, C.byref(swo))
            assert np.array_equal(vo, vc)
            assert swo.value == sw.value


def test_tc1_file_out_counts():
    """FILE.OUT of the reference's tc1 run: NELEM NPOIN NBOUN = 28304 14358
    410 and RORDER TORDER VNPNT = 40 40 1600 (tc/tc1/FILE.OUT:5,8)."""
    txt = open(os.path.join(GOLDEN, "tc1_FILE.OUT")).read().split()
    nums = [int(x) for x in txt if x.isdigit()]
    assert nums[-6:] == [28304, 14358, 410, 40, 40, 1600]


def test_gteqnf_and_getfor_bit_exact():
    from oracle_py import lib
    R = _ref()
    rng = np.random.default_rng(1)
    E = 257
    nd = rng.random(3 * E) * 2 + 0.1
    ux = rng.standard_normal(3 * E) * 0.3
    uy = rng.standard_normal(3 * E) * 0.3
    ps = rng.random(3 * E) + 0.5
    ps[::17] = 1e-25  # vacuum branch (src/gteqnf.f90:32)
    for (cx, cy) in [(0.3, -1.2), (2.5, 0.1)]:
        a, b = np.zeros(3 * E), np.zeros(3 * E)
        lib().ora_gteqnf(E, nd, ux, uy, ps, cx, cy, 6.022e26, a)
        R.L.ref_gteqnf(C.c_int(E), _ptr(nd), _ptr(ux), _ptr(uy), _ptr(ps), C.c_double(cx),
                       C.c_double(cy), C.c_double(6.022e26), _ptr(b))
        assert np.array_equal(a, b)
    case = cylinder_small()
    m = case.mesh
    psp = 1.0 + rng.random(m.npoin)
    lo, do = C.c_double(), C.c_double()
    lr, dr = C.c_double(), C.c_double()
    lib().ora_getfor(m.nboun, _ip(m.bsido), psp, np.ascontiguousarray(m.coord), C.byref(lo),
                     C.byref(do))
    R.L.ref_getfor(C.c_int(m.nboun), C.c_int(m.npoin), _ptr(_ip(m.bsido)), _ptr(psp),
                   _ptr(np.ascontiguousarray(m.coord)), C.byref(lr), C.byref(dr))
    assert (lo.value, do.value) == (lr.value, dr.value)


def _perturbed(o, seed=0):
    """A positive f with distinct values per ordinate so that a wrong
    ordinate index in a reflection shows up."""
    f = o.get_f()
    E, V, _ = f.shape
    rng = np.random.default_rng(seed)
    iv = np.arange(V)[None, :, None]
    f = f * (1.0 + 0.2 * np.sin(0.7 * iv + 0.37 * np.arange(E)[:, None, None])) \
        + 1e-3 * rng.random(f.shape)
    o.set_f(f)
    return o.get_f()


def _interior_sides(rhs, f, umean, iside, nx, ny, el, iv, ux, uy):
    """The interior branch of EDGFLXOPT (src/edgflxopt.f90:95-125), applied
    in side order on the reference's own arrays."""
    mm = ((2.0, 1.0), (1.0, 2.0))
    for s, row in enumerate(iside):
        ip1, ip2, iel, ier, inl1, inl2, inr1, inr2 = (int(x) for x in row)
        if ier == 0:
            continue
        scpr = ux * nx[s] + uy * ny[s]
        if scpr > 0.0:
            eu = (f[iel - 1, iv - 1, inl1 - 1] - umean[iel - 1],
                  f[iel - 1, iv - 1, inl2 - 1] - umean[iel - 1])
        else:
            eu = (f[ier - 1, iv - 1, inr1 - 1] - umean[ier - 1],
                  f[ier - 1, iv - 1, inr2 - 1] - umean[ier - 1])
        fx = [0.0 if abs(x) == 0.0 else ux * x for x in eu]
        fy = [0.0 if abs(x) == 0.0 else uy * x for x in eu]
        alen = (1.0 / 6.0) * el[s]
        fn = [nx[s] * fx[j] + ny[s] * fy[j] for j in range(2)]
        rh = [0.0, 0.0]
        for i in range(2):
            for j in range(2):
                rh[i] = rh[i] - fn[j] * mm[i][j] * alen
        rhs[ier - 1, inr1 - 1] -= rh[0]
        rhs[ier - 1, inr2 - 1] -= rh[1]
        rhs[iel - 1, inl1 - 1] += rh[0]
        rhs[iel - 1, inl2 - 1] += rh[1]


@pytest.mark.parametrize("make", [c1, cylinder_small])
def test_conmat_bit_exact(make):
    """CONMAT (src/conmat.f90:1-44) against the reference's own, on the
    reference's GEOME; the oracle context built with IMMAT = 0 holds the same
    matrices."""
    from oracle_py import OracleRun, lib
    case = make()
    case.prm["immat"] = 0
    R = _ref()
    E = case.mesh.nelem
    t = _ref_setup(R, case.mesh)
    geome = np.ascontiguousarray(t["geome"])
    ref = np.zeros(9 * E)
    R.L.ref_conmat(C.c_int(E), _ptr(geome), _ptr(ref))
    ora = np.zeros(9 * E)
    lib().ora_conmat(E, geome.ravel(), ora)
    assert np.array_equal(ora, ref)
    o = OracleRun(case)
    assert np.array_equal(o.cmmat().ravel(), ref)
    o.close()
    # the row sums are the lumped inverse mass 6/RJ (GETMAT's value)
    rows = ref.reshape(E, 3, 3).sum(axis=1)
    assert np.allclose(rows, t["mmat"], rtol=1e-13)


@pytest.mark.parametrize("make,lookup,immat", [(lobatto10_walls, 0, 1), (cylinder_lobatto10, 0, 1),
                                               (lobatto10_walls, 0, 0), (cylinder_lobatto10, 0, 0)])
def test_one_ordinate_stage_by_stage(make, lookup, immat):
    """Every ordinate of one step, in order, composed from the reference's
    GTEQNF, GETFLA+GETELC, GETBOU, GETINC+ADTHEM (or, with IMMAT = 0,
    CMMINC+ADTHEM on CONMAT's matrices) and INFLOW on the same arrays (so
    reflections read ordinates already advanced this step), against the
    oracle.  Grids where the reference's ring rule is defined (RORDER ==
    TORDER, first node -1)."""
    from oracle_py import OracleRun
    case = make()
    case.prm["alpha"] = 0.6  # both the diffuse and the specular part of GETBOU
    case.prm["immat"] = immat
    R = _ref()
    o = OracleRun(case, lookup=lookup)
    f = _perturbed(o).copy()
    dte = o.prepare_step()
    nf = o.node_fields()
    eta = np.ascontiguousarray(o.eta())
    t = o.tables()
    E, V = case.mesh.nelem, case.nvel
    iside = o.iside()
    vcord = np.ascontiguousarray(o.vcord())
    bsido = _ip(case.mesh.bsido)
    rsido = np.ascontiguousarray(t["rsido"])
    geome = np.ascontiguousarray(t["geome"])
    mmat = np.ascontiguousarray(t["mmat"])
    cmmat = np.zeros(9 * case.mesh.nelem)
    R.L.ref_conmat(C.c_int(case.mesh.nelem), _ptr(geome), _ptr(cmmat))
    mu = np.ascontiguousarray(nf["mu"])
    cinf = o.cinf()
    walls = 0
    for iv in range(1, V + 1):
        ux, uy = o.velocity(iv)
        st = o.debug_ordinate(iv, dte)
        nfo = np.zeros((E, 3))
        R.L.ref_gteqnf(C.c_int(E), _ptr(nf["disnd"]), _ptr(nf["disux"]), _ptr(nf["disuy"]),
                       _ptr(nf["disps"]), C.c_double(ux), C.c_double(uy),
                       C.c_double(case.prm["m"]), _ptr(nfo))
        assert np.array_equal(nfo.ravel(), st["nfo"]), iv
        umean, rhs = np.zeros(E), np.zeros((E, 3))
        R.L.ref_getelc(C.c_int(E), C.c_int(V), C.c_int(iv), _ptr(geome), _ptr(f),
                       C.c_double(ux), C.c_double(uy), C.c_double(dte), _ptr(mu), _ptr(nfo),
                       _ptr(umean), _ptr(rhs))
        assert np.array_equal(umean, st["umean"]), iv
        assert np.array_equal(rhs.ravel(), st["rhs_elc"]), iv
        for s, row in enumerate(iside):
            if row[3] == 0:
                btest = ux * rsido[:, 0] + uy * rsido[:, 1]
                walls += int((btest < 0).any())
                R.L.ref_getbou(C.c_int(case.mesh.nboun), C.c_int(E), _ptr(bsido),
                               C.c_int(int(row[2])), _ptr(rsido), C.c_int(int(row[0])),
                               _ptr(rhs), C.c_int(int(row[4])), C.c_int(int(row[5])),
                               C.c_double(ux), C.c_double(uy), C.c_double(case.prm["alpha"]),
                               _ptr(eta), C.c_int(V), C.c_int(iv), _ptr(f), _ptr(umean),
                               C.c_double(case.prm["rv"]), _ptr(vcord),
                               C.c_int(case.rorder), C.c_int(case.torder),
                               C.c_double(case.prm["rgas"]))
            else:
                _interior_sides(rhs, f, umean, iside[s:s + 1], t["nx"][s:s + 1],
                                t["ny"][s:s + 1], t["el"][s:s + 1], iv, ux, uy)
        assert np.array_equal(rhs.ravel(), st["rhs_edge"]), iv
        rhs = dte * rhs
        residual = np.zeros(V)
        if immat:
            R.L.ref_getinc_adthem(C.c_int(E), C.c_int(V), C.c_int(iv), _ptr(mmat), _ptr(rhs),
                                  _ptr(f), _ptr(nfo), _ptr(mu), C.c_double(dte),
                                  _ptr(residual))
        else:
            R.L.ref_cmminc_adthem(C.c_int(E), C.c_int(V), C.c_int(iv), _ptr(cmmat), _ptr(rhs),
                                  _ptr(f), _ptr(nfo), _ptr(mu), C.c_double(dte),
                                  _ptr(residual))
        assert residual[iv - 1] == st["residual"], iv
        R.L.ref_inflow(C.c_int(iv), C.c_double(ux), C.c_double(uy), C.c_int(E),
                       _ptr(_ip(case.mesh.intma)), _ptr(cinf), C.c_int(case.mesh.nboun),
                       _ptr(bsido), C.c_int(V), _ptr(f), C.c_double(case.prm["rgas"]),
                       C.c_double(case.prm["m"]))
        assert np.array_equal(f, o.get_f()), iv
    assert walls > 0  # reflections were exercised


def test_alotim2_bit_exact():
    from oracle_py import OracleRun
    R = _ref()
    case = cylinder_small()
    o = OracleRun(case)
    dte = o.prepare_step()
    mu = np.ascontiguousarray(o.node_fields()["mu"])
    dteb = C.c_double()
    R.L.ref_alotim2(C.c_int(case.mesh.nelem), _ptr(mu), C.byref(dteb))
    t = o.tables()
    assert dte == min(dteb.value, t["dtec"])


@pytest.mark.parametrize("make", [lobatto10_walls, cylinder_lobatto10])
def test_corrected_ring_rule_equals_reference_rule_where_defined(make):
    """On grids where the reference's ring lookup is defined, the corrected
    rule reads the same ordinates: whole steps agree bit for bit."""
    from oracle_py import OracleRun
    case = make()
    a, b = OracleRun(case, lookup=0), OracleRun(case, lookup=1)
    ra, rb = a.step(4), b.step(4)
    assert ra == rb
    assert np.array_equal(a.get_f(), b.get_f())


def test_reference_ring_rule_rejected_where_undefined():
    """RORDER != TORDER: the reference rule reads outside the ring table
    (src/getbou.f90:223-243); the oracle reports it instead of reading."""
    from oracle_py import OracleRun
    case = cylinder_small()
    o = OracleRun(case, lookup=0)
    with pytest.raises(RuntimeError):
        o.step(1)
