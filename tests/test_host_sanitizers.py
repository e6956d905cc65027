"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize builds the library's host setup (dg-bgk_amd/csrc/prep.cpp:
prepare, the storage order of a device context, the reflection table) and
the oracle (oracle/dgbgk_oracle.c) with -fsanitize=address,undefined into
two small drivers (g++/gcc, CPU only; any report aborts the driver).  Each
driver runs on the test decks, including a shuffled mesh, the reference
ring rule and an OpenMP-threaded oracle, and its outputs must equal, bit
for bit, what the normally built libraries give on the same input.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from small_cases import c1, cylinder_lobatto10, cylinder_small, lobatto10_walls, shuffled

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")

pytestmark = pytest.mark.skipif(
    not os.path.isdir(SAN) or shutil.which("g++") is None or shutil.which("make") is None,
    reason="sanitizer drivers or host compilers absent")

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    out = tmp_path_factory.mktemp("san")
    r = subprocess.run(["make", "-s", "-C", SAN, f"OUT={out}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def _write_case(d, case, params):
    m = case.mesh
    os.makedirs(d, exist_ok=True)
    np.array([m.nelem, m.npoin, m.nboun, case.rorder, case.torder], np.int32).tofile(
        os.path.join(d, "sizes.i32"))
    np.ascontiguousarray(m.intma, np.int32).tofile(os.path.join(d, "intma.i32"))
    np.ascontiguousarray(m.coord, np.float64).tofile(os.path.join(d, "coord.f64"))
    np.ascontiguousarray(m.bsido, np.int32).tofile(os.path.join(d, "bsido.i32"))
    np.ascontiguousarray(case.eta, np.float64).tofile(os.path.join(d, "eta.f64"))
    np.ascontiguousarray(case.w, np.float64).tofile(os.path.join(d, "w.f64"))
    with open(os.path.join(d, "params.bin"), "wb") as fh:
        fh.write(C.string_at(C.addressof(params), C.sizeof(params)))


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=900)
    assert r.returncode == 0, (r.returncode, r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr


def _read(d, name, dtype):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


def _shuffled_cylinder():
    return shuffled(cylinder_small(), seed=3)[0]


PREP_CASES = {"c1": (c1, 1), "cylinder_small": (cylinder_small, 1),
              "cylinder_shuffled": (_shuffled_cylinder, 1),
              "lobatto10_walls_reference_rule": (lobatto10_walls, 0),
              "cylinder_lobatto10_reference_rule": (cylinder_lobatto10, 0)}


@pytest.mark.parametrize("name", list(PREP_CASES))
def test_host_setup_under_sanitizers(drivers, tmp_path, name):
    from dgbgk import _lib
    from dgbgk.solver import Setup
    make, lookup = PREP_CASES[name]
    case = make()
    _write_case(tmp_path / "in", case, _lib.make_params(case.prm, lookup=lookup))
    os.makedirs(tmp_path / "out")
    _run([str(drivers / "prep_check"), str(tmp_path / "in"), str(tmp_path / "out")])
    out = tmp_path / "out"
    s = Setup(case, lookup=lookup)
    t, r = s.tables(), s.reflections()
    ntiles = s.ntiles
    s.close()
    assert np.array_equal(_read(out, "iside.i32", np.int32), t["iside"].ravel())
    for k in ("nx", "ny", "el", "geome", "relen", "mmat", "rsido", "vcord"):
        assert np.array_equal(_read(out, f"{k}.f64", np.float64), t[k].ravel()), k
    assert np.array_equal(_read(out, "scalars.f64", np.float64), [t["sumweight"], t["dtec"]])
    assert np.array_equal(_read(out, "wall_sides.i32", np.int32) + 1, r["side"])
    assert np.array_equal(_read(out, "refl_i1.i32", np.int32) + 1, r["i1"].ravel())
    assert np.array_equal(_read(out, "refl_i2.i32", np.int32) + 1, r["i2"].ravel())
    assert np.array_equal(_read(out, "refl_zd1.f64", np.float64), r["zd1"].ravel())
    assert np.array_equal(_read(out, "refl_zd2.f64", np.float64), r["zd2"].ravel())
    assert _read(out, "ntiles.i32", np.int32)[0] == ntiles
    eorder = _read(out, "eorder.i32", np.int32)
    if eorder.size:
        assert np.array_equal(np.sort(eorder), np.arange(case.mesh.nelem))


ORACLE_CASES = {"c1": (c1, 1, 3), "cylinder_small": (cylinder_small, 1, 2),
                "lobatto10_walls_reference_rule": (lobatto10_walls, 0, 3)}


@pytest.mark.parametrize("name", list(ORACLE_CASES))
def test_oracle_under_sanitizers(drivers, tmp_path, name):
    from oracle_py import OracleRun, OraStepInfo, make_params
    make, lookup, steps = ORACLE_CASES[name]
    case = make()
    _write_case(tmp_path / "in", case, make_params(case.prm, lookup))
    os.makedirs(tmp_path / "out")
    _run([str(drivers / "oracle_check"), str(tmp_path / "in"), str(tmp_path / "out"),
          str(steps), "4"])
    out = tmp_path / "out"
    o = OracleRun(case, lookup=lookup)
    recs = o.step(steps)
    f = o.get_f()
    o.close()
    nrec = len(OraStepInfo._fields_)
    want = np.array([[r[k] for k, _ in OraStepInfo._fields_] for r in recs])
    for tag in ("serial", "threads"):
        info = _read(out, f"info_{tag}.f64", np.float64).reshape(steps, nrec)
        assert np.array_equal(info, want), tag
        assert np.array_equal(_read(out, f"f_{tag}.f64", np.float64), f.ravel()), tag
