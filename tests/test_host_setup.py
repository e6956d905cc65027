"""The product library's host side, without a GPU: it loads, exports every
entry point of include/dgbgk.h, rebuilds the reference's setup tables bit
for bit, and turns the reference's STOP conditions into status codes."""
import os
import re

import numpy as np
import pytest

from small_cases import c1, cylinder_lobatto10, cylinder_small, lobatto10_walls, shuffled

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from dgbgk import _lib
    L = _lib.lib()
    hdr = open(os.path.join(ROOT, "include", "dgbgk.h")).read()
    declared = set(re.findall(r"\b(dgbgk_[a-z_]+)\s*\(", hdr))
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def _cylinder_shuffled():
    return shuffled(cylinder_small(), seed=3)[0]


@pytest.mark.parametrize("make,lookup", [(c1, 1), (cylinder_small, 1), (lobatto10_walls, 0),
                                         (cylinder_lobatto10, 0), (_cylinder_shuffled, 1)])
def test_setup_tables_match_oracle_bit_for_bit(make, lookup):
    """The reference's tables in the input element numbering, also where the
    library keeps the elements in another order."""
    from dgbgk.solver import Setup
    from oracle_py import OracleRun
    case = make()
    s = Setup(case, lookup=lookup)
    o = OracleRun(case, lookup=lookup)
    t, ot = s.tables(), o.tables()
    assert np.array_equal(t["iside"], o.iside())
    assert np.array_equal(t["vcord"], o.vcord())
    assert t["sumweight"] == o.sumweight()
    for k in ("nx", "ny", "el", "geome", "relen", "mmat", "rsido", "dtec"):
        assert np.array_equal(t[k], ot[k]), k
    assert s.ring_separable


def test_tiles_cover_every_element_once():
    """The update kernel's tiles: contiguous owned ranges covering the mesh,
    owned + halo within the 256 LDS slots (setup fails with DGBGK_ERR_TILE
    otherwise)."""
    from dgbgk.solver import Setup
    for case in (c1(), cylinder_small(elems=6000)):
        s = Setup(case)
        assert s.ntiles >= (case.mesh.nelem + 255) // 256
        # halo slots take at most a third of a tile on these meshes
        assert s.ntiles <= (case.mesh.nelem + 169) // 170


def test_tiles_do_not_depend_on_the_input_numbering(monkeypatch):
    """The library keeps elements in Hilbert order of their centroids, so a
    mesh in random element order gets the same compact tiles as the mesh
    generator's own (Hilbert) order; kept in the input order
    (DGBGK_REORDER=0) its tiles would hold mostly halo elements."""
    from dgbgk.solver import Setup
    case = cylinder_small(elems=6000)
    sh, _ = shuffled(case, seed=1)
    n0 = Setup(case).ntiles
    assert Setup(sh).ntiles == n0
    monkeypatch.setenv("DGBGK_REORDER", "0")
    assert Setup(case).ntiles == n0          # the generator's order is kept as is
    assert Setup(sh).ntiles > 2 * n0


def _err(case, **kw):
    from dgbgk import _lib
    from dgbgk.solver import Setup
    with pytest.raises(_lib.DgbgkError) as ei:
        Setup(case, **kw)
    return ei.value.code


def test_clockwise_element_is_rejected():
    case = c1(n=4)
    case.mesh.intma[3] = case.mesh.intma[3][[0, 2, 1]]
    assert _err(case) == -7


def test_boundary_side_missing_from_bsido():
    """GETBOU's linear search runs past NBOUN when a boundary side of the mesh
    is not in BSIDO (src/getbou.f90:48-53)."""
    case = c1(n=4)
    case.mesh.bsido = case.mesh.bsido[1:].copy()
    assert _err(case) == -3


def test_sumbottom_stop():
    """A type-4 wall at Tw = 0 gives SUMBOTTOM < 0.1 (src/flucon2.f90:227-231)."""
    case = lobatto10_walls(n=4)
    case.mesh.bsido[case.mesh.bsido[:, 3] == 4, 4] = 0
    assert _err(case) == -5


def test_reference_ring_rule_outside_its_domain():
    case = cylinder_small(elems=600)
    assert _err(case, lookup=0) == -6


def test_mass_matrix_choice():
    """IMMAT = 0 (the consistent-mass branch) builds; other values are
    rejected; with IMMAT = 0 the lumped MMAT table stays zero, as in the
    reference, which then fills CMMAT instead (src/dg_lin_convec.f90:113-117)."""
    from dgbgk.solver import Setup
    from oracle_py import OracleRun
    case = c1(n=4)
    case.prm["immat"] = 2
    assert _err(case) == -1
    case.prm["immat"] = 0
    t = Setup(case).tables()
    o = OracleRun(case)
    assert not t["mmat"].any()
    assert np.array_equal(t["mmat"], o.tables()["mmat"])
    assert o.cmmat().any()
    o.close()


def test_inflow_without_inflow_state():
    case = c1(n=4)
    case.prm["inf"] = 0
    assert _err(case) == -1


def test_point_id_out_of_range():
    case = c1(n=4)
    case.mesh.intma[0, 0] = case.mesh.npoin + 5
    assert _err(case) == -7
