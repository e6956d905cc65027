"""The FP32-bulk store near a steady state.  Rounding f + d to nearest drops
every increment below half an ulp, so a run stalls at a state biased by
about ulp / (decay rate per step), a bias that grows as the time step
shrinks.  The library's FP32 store rounds with a dither of +-1/2 ulp
instead (unbiased), and the steady drag then agrees with the FP64 run to
rounding level.  Measured on this case (1,679 triangles, 8 x 16
ordinates, CFL factor 0.25, 80,000 steps): drag relative to FP64 5.5e-5
with the plain store, 8e-7 with the dithered one; step change of the
dithered FP32 run 2.6e-7 of step 1's."""
import pytest

pytestmark = pytest.mark.gpu


def _steady_drag(case, precision, nsteps):
    from dgbgk.solver import Solver
    s = Solver(case, precision=precision)
    first = None
    done = 0
    while done < nsteps:
        r = s.step(2000)
        done += 2000
        if first is None:
            first = r[0]["change"]
    _, drag = s.forces()
    s.close()
    return drag, r[-1]["change"] / first


def test_dithered_fp32_store_has_no_steady_state_bias(monkeypatch):
    from dgbgk import _lib, warmstart
    mesh = warmstart.mesh_for(1200)
    case = warmstart.cylinder_case(mesh, warmstart.Sample(0.8, 0.15), 8, 16)
    case.prm["csafm"] = 0.25
    n = 80000
    print(f"case cylinder E={mesh.nelem} steady drag, {n} steps x 3 runs", flush=True)
    d64, c64 = _steady_drag(case, _lib.DGBGK_FP64, n)
    d32, c32 = _steady_drag(case, _lib.DGBGK_FP32, n)
    monkeypatch.setenv("DGBGK_FP32_DITHER", "0")
    d32p, _ = _steady_drag(case, _lib.DGBGK_FP32, n)
    dith = abs(d32 - d64) / abs(d64)
    plain = abs(d32p - d64) / abs(d64)
    print(f"drag rel to FP64: dithered {dith:.2e}, plain {plain:.2e}; change/first "
          f"FP64 {c64:.1e}, FP32 {c32:.1e}", flush=True)
    assert c64 < 1e-12          # the FP64 run is steady
    assert c32 < 1e-6           # the dithered run's change floor is below 1e-6
    assert dith < 5e-6
    assert plain > 1e-5         # the bias the dither removes
