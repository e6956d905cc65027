"""EDGFLXOPT's interior-side flux against direct quadrature of its integral.

The reference's interior loop (src/edgflxopt.f90:95-125) adds to each
element's right-hand side the edge term of the Taylor-Galerkin scheme,

    R[e, i] += - integral over each interior edge of N_i^e (c . n_e) g_up ds,

where n_e is e's outward unit normal, g = f - U (the element's P1 values
less its predictor U), and g_up is the trace of the element that c . n
points away from (on a tie, the right element).  This test evaluates that
integral independently: edges found from shared point pairs (not from
ISIDE), normals from the coordinates, P1 bases from barycentric
coordinates (not from GEOME), and two-point Gauss-Legendre quadrature
along each edge, which is exact for the quadratic integrand.  The oracle's
edge contribution is its right-hand side after EDGFLXOPT less the one after
GETELC, one ordinate at a time, with mu = 0 (hard-sphere diameter 0) and
every boundary side of inflow type (no boundary flux).  Its f is linear
per element and discontinuous between elements, on a jittered 8-triangle
mesh, at three ordinates.  The planted misreadings (downwind values,
ISIDE(5:6) swapped) must fail the same comparison.
"""
import numpy as np
import pytest

GL2 = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))


def _case():
    from dgbgk import cases, meshgen
    m = meshgen.unit_square(2, left=1, right=1, bottom=1, top=1)
    rng = np.random.default_rng(7)
    interior = (m.coord[:, 0] > 0) & (m.coord[:, 0] < 1) & (m.coord[:, 1] > 0) & \
        (m.coord[:, 1] < 1)
    m.coord[interior] += rng.uniform(-0.12, 0.12, size=(interior.sum(), 2))
    prm = cases.base_params(u0=0.0, kn=1.0, length=1.0, rv=2.0, alpha=1.0)
    prm["d"] = 0.0                                  # sigma = 0, so mu = 0
    # one speed ring (r = rv/2 = 1) x 3 angles: theta = -60, 60, 180 degrees
    return cases.Case("edge_quadrature", m, np.array([0.0]), np.array([2.0]), 3, prm)


def _bary(tri_xy, x):
    """P1 basis values of a triangle (3x2 vertex array) at point x."""
    T = np.array([[tri_xy[0, 0], tri_xy[1, 0], tri_xy[2, 0]],
                  [tri_xy[0, 1], tri_xy[1, 1], tri_xy[2, 1]],
                  [1.0, 1.0, 1.0]])
    return np.linalg.solve(T, np.array([x[0], x[1], 1.0]))


def _expected_edge_rhs(case, g, c):
    """-sum over interior edges of int N_i (c.n_e) g_up ds per element node."""
    intma = case.mesh.intma - 1
    xy = case.mesh.coord
    E = len(intma)
    owner = {}
    for e in range(E):
        for k in range(3):
            p, q = intma[e, k], intma[e, (k + 1) % 3]
            owner.setdefault(frozenset((p, q)), []).append((e, p, q))
    out = np.zeros((E, 3))
    for pair in owner.values():
        if len(pair) != 2:
            continue                                # boundary side (inflow: no flux)
        (e1, p, q), (e2, _, _) = pair
        P, Q = xy[p], xy[q]
        ln = np.hypot(*(Q - P))
        n1 = np.array([Q[1] - P[1], -(Q[0] - P[0])]) / ln   # outward for e1 (CCW)
        cn1 = float(c @ n1)
        # upwind: the element that c points out of (no ties c.n = 0 occur on
        # this jittered mesh at these angles)
        assert abs(cn1) > 1e-9
        up = e1 if cn1 > 0 else e2
        for e, sgn in ((e1, 1.0), (e2, -1.0)):
            for s, w in zip(GL2, (0.5, 0.5)):
                x = P + s * (Q - P)
                gu = _bary(xy[intma[up]], x) @ g[up]
                Ne = _bary(xy[intma[e]], x)
                out[e] -= w * ln * Ne * (sgn * cn1) * gu
    return out


def _oracle_edge_rhs(case, f, mutant=0):
    from oracle_py import OracleRun
    o = OracleRun(case, mutant=mutant)
    o.set_f(f)
    dte = o.prepare_step()
    assert np.all(o.node_fields()["mu"] == 0.0)
    res = []
    for iv in range(1, case.nvel + 1):
        d = o.debug_ordinate(iv, dte)
        res.append(((d["rhs_edge"] - d["rhs_elc"]).reshape(-1, 3), d["umean"].copy(),
                    np.array(o.velocity(iv))))
    o.close()
    return res


@pytest.mark.parametrize("mutant", [0, 1, 2], ids=["oracle", "downwind", "swapped_nodes"])
def test_interior_edge_flux_is_the_quadrature_of_its_integral(mutant):
    case = _case()
    E, V = case.mesh.nelem, case.nvel
    rng = np.random.default_rng(3)
    f = rng.uniform(0.5, 2.0, size=(E, V, 3))      # linear per element, discontinuous
    worst = 0.0
    for iv, (rhs, umean, c) in enumerate(_oracle_edge_rhs(case, f, mutant)):
        g = f[:, iv, :] - umean[:, None]
        want = _expected_edge_rhs(case, g, c)
        worst = max(worst, float(np.max(np.abs(rhs - want)) / np.max(np.abs(want))))
    if mutant == 0:
        assert worst < 1e-12, worst
    else:
        assert worst > 1e-2, worst
