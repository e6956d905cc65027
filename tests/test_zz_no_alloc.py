"""The allocation-free statistics entry point (dgbgk_f_stats_noalloc),
called through the C ABI, against dgbgk_f_stats: two calls in a row on
different contexts, the second with a smaller f, in both precisions, and
the size limit of its fixed scratch."""
import pytest

from small_cases import c1, cylinder_small

pytestmark = pytest.mark.gpu


def test_f_stats_noalloc_matches_f_stats():
    from dgbgk import _lib
    from dgbgk.solver import Solver
    runs = ((cylinder_small(), _lib.DGBGK_FP64, 7), (c1(n=8), _lib.DGBGK_FP32, 3))
    for case, prec, n in runs:
        print(f"no-alloc stats: {case.name} E={case.mesh.nelem} V={case.nvel} "
              f"precision {prec}", flush=True)
        with Solver(case, precision=prec) as s:
            s.step(n)
            a = s.f_stats()
            b = s.f_stats_noalloc()
            assert a == b, (a, b)
            assert a["nonfinite"] == 0 and a["min"] >= 0.0 < a["max"]
            b2 = s.f_stats_noalloc()   # the scratch keeps the last call's contents
            assert b2 == a


def test_f_stats_noalloc_rejects_more_ordinates_than_its_scratch():
    from dgbgk import _lib, cases, meshgen
    from dgbgk.solver import Solver
    eta, w = cases.gauss_legendre_speeds(65)
    case = cases.Case("wide_grid", meshgen.unit_square(2, left=2, right=2, bottom=2, top=2),
                      eta, w, 64, cases.base_params(u0=0.1, kn=0.1, length=1.0, rv=5.0,
                                                    alpha=1.0))
    print(f"no-alloc stats: V={case.nvel} (over the limit)", flush=True)
    with Solver(case) as s:
        with pytest.raises(_lib.DgbgkError) as ei:
            s.f_stats_noalloc()
        assert ei.value.code == -1   # DGBGK_ERR_ARG
        assert s.f_stats()["nonfinite"] == 0
