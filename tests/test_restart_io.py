"""Restart and GiD mesh files (no GPU): the fixed writer/reader pair
round-trips f exactly, rejects a file of another mesh or velocity grid,
and keeps the reference writer's line layout."""
import numpy as np
import pytest

from dgbgk import io
from small_cases import c1


def test_restart_round_trip_is_exact(tmp_path):
    case = c1(n=4)
    E, V, P = case.mesh.nelem, case.nvel, case.mesh.npoin
    rng = np.random.default_rng(5)
    f = rng.random((E, V, 3)) * 10.0 ** rng.integers(-30, 3, (E, V, 3))
    path = str(tmp_path / "RESTART.RES")
    io.write_restart(path, f, P, case.eta, case.torder, case.prm["rv"])
    g = io.read_restart(path, E, V, P)
    assert np.array_equal(f, g)
    lines = open(path).read().splitlines()
    assert lines[0] == "NPOIN     VNPNT"
    assert [int(t) for t in lines[1].split()] == [P, V]
    assert len(lines) == 2 + E * V
    # element id, ordinate velocity, three nodal values; ordinate 1 first
    first = lines[2].split()
    assert int(first[0]) == 1 and len(first) == 6
    cx, cy = io._ordinate_xy(case.eta, case.torder, case.prm["rv"])
    assert abs(float(first[1]) - cx[0]) <= 1e-6 * abs(cx[0]) + 1e-12
    assert abs(float(first[2]) - cy[0]) <= 1e-6 * max(abs(cy[0]), 1.0)
    with pytest.raises(ValueError):
        io.read_restart(path, E, V, P + 1)


def test_ordinate_velocities_match_the_oracle_grid():
    from oracle_py import OracleRun
    case = c1(n=4)
    o = OracleRun(case)
    vc = o.vcord()  # (V, 3): eta, zeta, weight of every ordinate
    cx, cy = io._ordinate_xy(case.eta, case.torder, case.prm["rv"])
    rt = vc[:, 0] * (case.prm["rv"] / 2) + case.prm["rv"] / 2
    assert np.allclose(cx, rt * np.cos(vc[:, 1] * 3.1416), rtol=0, atol=1e-12)
    assert np.allclose(cy, rt * np.sin(vc[:, 1] * 3.1416), rtol=0, atol=1e-12)


def test_gid_mesh_round_trip(tmp_path):
    case = c1(n=6)
    path = str(tmp_path / "GIDMESH.RES")
    io.write_gid_mesh(path, case.mesh)
    coord, intma = io.read_gid_mesh(path)
    assert np.array_equal(intma, case.mesh.intma)
    assert np.allclose(coord, case.mesh.coord, atol=6e-6)
    txt = open(path).read().splitlines()
    assert txt[0] == "mesh dimension = 2 elemtype triangle nnode = 3"
    assert txt[-1] == "end elements"
