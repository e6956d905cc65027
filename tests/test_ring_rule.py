"""The wall reflection table of GETBOU (src/getbou.f90:190-250) as the
library's host setup builds it (dgbgk_setup_create, dgbgk_setup_reflections),
checked from geometry alone on the velocity grids of the workloads: no
table, oracle or reference rule is consulted.

An ordinate c that enters the domain through a wall side with outward
normal n (c.n < 0) takes its value from the reflected velocity
c' = c - 2 (c.n) n.  The table names the two ordinates c' is interpolated
between.  Whatever the rule that finds them, on a polar grid (speed
(RV/2)(ETA+1), angle PI*ZETA with the reference's PI = 3.1416,
src/getbou.f90:41) it must satisfy:
- only entering ordinates reflect;
- both ordinates sit on the speed ring of radius |c'| = |c|;
- their ZETA bracket contains the angle of c' (periodically, where the
  bracket wraps from the ring's last angle to its first);
- zd1 and zd2 are the bracket's width and the offset of c' in it (the
  reference's wrap branch measures both from ZETA = -1).
GETBOU takes the angle of c' as PI - acos|c'x/|c'|| in the left half
plane, so with PI = 3.1416 its ZETA differs from the true angle / PI by up
to (3.1416 - pi) / PI = 2.3e-6; brackets and offsets are checked to 1e-5
in ZETA (a node spacing is at least 1/32 on these grids).
The checker is itself checked: tables with a planted mistake (the
neighbouring ring, the two ordinates swapped, a flipped normal) fail it.
"""
import numpy as np
import pytest

from small_cases import c1, cylinder_lobatto10, lobatto10_walls


PI = 3.1416
ZTOL = 1e-5


def _grid(vcord, rv):
    eta, zeta = vcord[:, 0], vcord[:, 1]
    speed = 0.5 * rv * (eta + 1.0)
    c = np.stack([speed * np.cos(PI * zeta), speed * np.sin(PI * zeta)], axis=1)
    return c, speed, zeta


def violations(vcord, rv, normals, refl, flip_normal=False):
    """Every way the reflection table breaks the properties above."""
    c, speed, zeta = _grid(vcord, rv)
    bad = []
    tol = 1e-12
    for k, n in enumerate(normals):
        if flip_normal:
            n = -n
        dot = c @ n
        i1, i2 = refl["i1"][k], refl["i2"][k]
        enter = dot < 0.0
        if np.any((i1 > 0) != enter) or np.any((i2 > 0) != enter):
            bad.append((k, "reflects an ordinate that does not enter"))
            continue
        for v in np.nonzero(enter)[0]:
            a, b = i1[v] - 1, i2[v] - 1
            cr = c[v] - 2.0 * dot[v] * n
            r = np.hypot(*cr)
            if abs(r - speed[v]) > tol * speed[v]:
                bad.append((k, v, "|c'| != |c|"))
            if abs(speed[a] - r) > tol * r or abs(speed[b] - r) > tol * r:
                bad.append((k, v, "bracket off the ring of radius |c'|"))
                continue
            zr = np.arctan2(cr[1], cr[0]) / PI
            za, zb = zeta[a], zeta[b]
            if za <= zb:                      # interior bracket
                if not (za - ZTOL <= zr <= zb + ZTOL):
                    bad.append((k, v, "angle outside its bracket"))
                elif (abs(refl["zd1"][k, v] - (zb - za)) > tol or
                      abs(refl["zd2"][k, v] - (zr - za)) > ZTOL):
                    bad.append((k, v, "bracket width or offset"))
            else:                             # wraps from the last angle to the first
                ring = speed == speed[a]
                if not (za == zeta[ring].max() and zb == zeta[ring].min()):
                    bad.append((k, v, "wrap bracket is not (last, first) of the ring"))
                elif zb + ZTOL < zr < za - ZTOL:
                    bad.append((k, v, "wrapped although an interior bracket holds the angle"))
                elif (abs(refl["zd1"][k, v] - (zb + 1.0)) > tol or
                      abs(refl["zd2"][k, v] - (zr + 1.0)) > ZTOL):
                    bad.append((k, v, "wrap width or offset"))
    return bad


def _cases():
    from dgbgk import cases
    return {
        "unit_square_8x16": (c1, 1),
        "cylinder_16x32": (lambda: cases.cylinder_flow(target_elems=1500), 1),
        "particle_32x64": (lambda: cases.particle_drag(target_elems=3000), 1),
        "cylinder_gl_8x16": (lambda: cases.cylinder_flow(target_elems=1500, rorder=8,
                                                         torder=16), 1),
        "lobatto10_walls_reference_rule": (lobatto10_walls, 0),
        "cylinder_lobatto10_reference_rule": (cylinder_lobatto10, 0),
    }


def _table(name):
    from dgbgk.solver import Setup
    make, lookup = _cases()[name]
    case = make()
    s = Setup(case, lookup=lookup)
    t, r = s.tables(), s.reflections()
    s.close()
    normals = t["rsido"][r["side"] - 1, :2]
    return case, t["vcord"], normals, r


@pytest.mark.parametrize("name", list(_cases()))
def test_reflection_table_from_geometry(name):
    case, vcord, normals, refl = _table(name)
    assert len(normals) > 0
    assert np.allclose(np.hypot(normals[:, 0], normals[:, 1]), 1.0, atol=1e-12)
    nref = int(np.count_nonzero(refl["i1"]))
    assert nref > 0
    bad = violations(vcord, case.prm["rv"], normals, refl)
    assert not bad, bad[:10]


def test_checker_rejects_planted_mistakes():
    case, vcord, normals, refl = _table("cylinder_16x32")
    T = case.torder
    V = len(vcord)

    def planted(**kw):
        r = {k: v.copy() for k, v in refl.items()}
        for k, f in kw.items():
            r[k] = f(r)
        return r

    on = refl["i1"] > 0
    # the neighbouring ring (what the reference rule's index gives off its grids)
    ring_up = planted(i1=lambda r: np.where(on, (r["i1"] - 1 + T) % V + 1, 0),
                      i2=lambda r: np.where(on, (r["i2"] - 1 + T) % V + 1, 0))
    swapped = planted(i1=lambda r: r["i2"], i2=lambda r: r["i1"])
    for label, r, flip in (("ring", ring_up, False), ("swap", swapped, False),
                           ("normal", refl, True)):
        assert violations(vcord, case.prm["rv"], normals, r, flip_normal=flip), label
