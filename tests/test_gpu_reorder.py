"""Element order on the device.  The library keeps elements in Hilbert
order of their centroids, whatever the input numbering, and maps f and the
node fields back to the input numbering at the C ABI.  The update of an
element reads only that element and its neighbours, and every sum over
ordinates runs in the same order, so a mesh and the same mesh with its
elements shuffled give bit-identical f, node fields, residuals and time
steps (the point averages keep the input numbering's MACROS order, so they
agree to rounding).  Parity against the oracle, which runs in the input
numbering, holds as for any other mesh."""
import numpy as np
import pytest

from small_cases import c1, cylinder_small, rel_l2, rel_l2_vec, shuffled

pytestmark = pytest.mark.gpu


def _run(case, precision, nsteps):
    from dgbgk.solver import Solver
    s = Solver(case, precision=precision)
    recs = s.step(nsteps)
    out = dict(f=s.get_f(), nf=s.node_fields(), pf=s.point_fields(), recs=recs,
               tiles=s.launch_config()["ntiles"], drag=s.forces()[1])
    s.close()
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_shuffled_mesh_runs_bit_identical(prec):
    case = cylinder_small()
    sh, perm = shuffled(case, seed=7)
    n = 130  # two captured 64-step graphs and two single steps
    print(f"case cylinder E={case.mesh.nelem} shuffled, precision {prec}, {n} steps",
          flush=True)
    a, b = _run(case, prec, n), _run(sh, prec, n)
    assert a["tiles"] == b["tiles"]
    assert np.array_equal(b["f"], a["f"][perm])
    for k in ("disnd", "disux", "disuy", "disps", "mu"):
        assert np.array_equal(b["nf"][k], a["nf"][k][perm]), k
    for ra, rb in zip(a["recs"], b["recs"]):
        for k in ("sumres", "dte", "change"):
            assert ra[k] == rb[k], k
        assert abs(ra["mass"] - rb["mass"]) <= 1e-13 * abs(ra["mass"])
    for k in ("nd", "rho", "ps", "temp"):
        assert rel_l2(b["pf"][k], a["pf"][k]) < 1e-14, k
    assert abs(b["drag"] - a["drag"]) <= 1e-12 * abs(a["drag"])


def test_shuffled_mesh_matches_oracle_fp64():
    from dgbgk.solver import Solver
    from oracle_py import OracleRun
    case, _ = shuffled(cylinder_small(), seed=11)
    print(f"case cylinder E={case.mesh.nelem} shuffled vs oracle, 20 steps", flush=True)
    s, o = Solver(case), OracleRun(case)
    rs, ro = s.step(20), o.step(20)
    for a, b in zip(rs, ro):
        assert abs(a["dte"] - b["dte"]) <= 1e-10 * abs(b["dte"])
        assert abs(a["mass"] - b["mass"]) <= 1e-12 * abs(b["mass"])
    assert rel_l2(s.get_f(), o.get_f()) < 1e-10
    nf, of = s.node_fields(), o.node_fields()
    for k in ("disnd", "disps"):
        assert rel_l2(nf[k], of[k]) < 1e-10, k
    assert rel_l2_vec(nf["disux"], nf["disuy"], of["disux"], of["disuy"]) < 1e-10
    pf, op = s.point_fields(), o.point_fields()
    for k in ("nd", "rho", "ps", "temp"):
        assert rel_l2(pf[k], op[k]) < 1e-10, k
    assert rel_l2(s.wall_eta(), o.eta()) < 1e-10


@pytest.mark.parametrize("prec", [0, 1])
def test_set_f_get_f_round_trip_in_input_numbering(prec):
    """set_f/get_f carry DISNF(3,V,E) in the input numbering: a state put in
    comes back unchanged (FP64) or rounded once to FP32, on a structured
    mesh (not in Hilbert order) and on a shuffled one; the node fields of one
    state on a mesh and on its shuffled copy agree under the shuffle."""
    from dgbgk.solver import Solver
    rng = np.random.default_rng(5)
    for case in (c1(n=16), cylinder_small()):
        sh, perm = shuffled(case, seed=2)
        print(f"case {case.name} E={case.mesh.nelem} set_f/get_f precision {prec}", flush=True)
        f = rng.uniform(0.5, 2.0, (case.mesh.nelem, case.nvel, 3))
        expect = f if prec == 0 else f.astype(np.float32).astype(np.float64)
        a, b = Solver(case, precision=prec), Solver(sh, precision=prec)
        a.set_f(f)
        b.set_f(f[perm])
        assert np.array_equal(a.get_f(), expect)
        assert np.array_equal(b.get_f(), expect[perm])
        na, nb = a.node_fields(), b.node_fields()
        for k in ("disnd", "disux", "disuy", "disps", "mu"):
            assert np.array_equal(nb[k], na[k][perm]), k
        a.close()
        b.close()


def test_context_setup_tables_are_the_input_meshs():
    """dgbgk_setup_tables on a context whose elements the library renumbered
    gives the reference's tables of the input mesh, bit for bit."""
    from dgbgk.solver import Solver
    from oracle_py import OracleRun
    case, _ = shuffled(cylinder_small(), seed=4)
    print(f"case cylinder E={case.mesh.nelem} shuffled, context setup tables", flush=True)
    s, o = Solver(case), OracleRun(case)
    t, ot = s.tables(), o.tables()
    assert np.array_equal(t["iside"], o.iside())
    assert np.array_equal(t["vcord"], o.vcord())
    assert t["sumweight"] == o.sumweight()
    for k in ("nx", "ny", "el", "geome", "relen", "mmat", "rsido", "dtec"):
        assert np.array_equal(t[k], ot[k]), k
    s.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_border_grouping_is_bit_identical(prec, monkeypatch):
    """Grouping tile border elements changes only where elements are stored
    inside the same tiles: f and the node fields are bit-identical with it
    and without it (DGBGK_BORDER_ORDER=0), in both forms (the FP32 store's
    dither phase belongs to the element, not to its storage position)."""
    case = cylinder_small()
    print(f"case cylinder E={case.mesh.nelem} border grouping on/off, precision {prec}",
          flush=True)
    a = _run(case, prec, 70)
    monkeypatch.setenv("DGBGK_BORDER_ORDER", "0")
    b = _run(case, prec, 70)
    assert a["tiles"] == b["tiles"]
    assert np.array_equal(a["f"], b["f"])
    for k in ("disnd", "disux", "disuy", "disps", "mu"):
        assert np.array_equal(a["nf"][k], b["nf"][k]), k
    for ra, rb in zip(a["recs"], b["recs"]):
        for k in ("sumres", "dte", "change"):
            assert ra[k] == rb[k], k
