"""The threaded oracle step (the all-core CPU baseline) equals the serial
restatement bit for bit: f, node and point fields and every step record,
on decks with inflow, outflow, diffuse walls and in-place specular
reflections read through the reference's ring rule."""
import numpy as np
import pytest

from small_cases import c1, cylinder_small, lobatto10_walls


@pytest.mark.parametrize("make,lookup", [(c1, 1), (cylinder_small, 1), (lobatto10_walls, 0)],
                         ids=["c1", "cylinder_small", "lobatto10_walls"])
def test_threaded_step_is_bit_identical(make, lookup):
    from oracle_py import OracleRun
    case = make()
    a, b = OracleRun(case, lookup=lookup), OracleRun(case, lookup=lookup)
    b.set_threads(4)
    ra, rb = a.step(6), b.step(6)
    assert ra == rb
    assert np.array_equal(a.get_f(), b.get_f())
    for k, v in a.node_fields().items():
        assert np.array_equal(v, b.node_fields()[k]), k
    for k, v in a.point_fields().items():
        assert np.array_equal(v, b.point_fields()[k]), k
