"""Closed-form and invariant pins of the oracle routines that the compiled
reference cannot run here (they call MPI): MACROS, COMPMU, FLUCON2 and the
interior-side loop of EDGFLXOPT.

A discrete Maxwellian on a fine grid must return its own density, velocity
and pressure; the collision frequency must equal sqrt(2)*sigma*n times the
2-D Maxwellian mean speed sqrt(pi)/(2 beta); the wall parameter ETA of a
wall Maxwellian must equal its outgoing number flux n/(2 sqrt(pi) beta)
over the reference's analytic denominator.  The reference evaluates angles
with PI = 3.1416, which biases angular sums by ~2e-6, so these hold to 1e-5
(FLUCON2's half-space sum has a kink at c.n = 0: 3e-3 at 40 angles).
The edge loop is pinned by free-stream preservation and by the pairwise
cancellation of interior fluxes.
"""
import math
import os

import numpy as np
import pytest

from small_cases import GOLDEN


def _fine_case(mesh, rv=8.0, alpha=1.0, kn=0.5):
    from dgbgk import cases, io
    eta, w = io.read_lobatto(os.path.join(GOLDEN, "Lobatto40.txt"))
    prm = cases.base_params(u0=0.0, kn=kn, length=1.0, rv=rv, alpha=alpha)
    return cases.Case("fine", mesh, eta, w, 40, prm)


def _maxwellian(o, n, u, v, T, rgas=1.0):
    """f = n beta^2/pi exp(-beta^2 |c-u|^2), beta^2 = 1/(2RT), on the
    ADVNCE velocity of each ordinate (true pi: this is the physics)."""
    E, V = o.E, o.V
    b2 = 1.0 / (2 * rgas * T)
    c = np.array([o.velocity(iv) for iv in range(1, V + 1)])
    g = n * b2 / math.pi * np.exp(-b2 * ((c[:, 0] - u) ** 2 + (c[:, 1] - v) ** 2))
    return np.broadcast_to(g[None, :, None], (E, V, 3)).copy()


def test_macros_compmu_of_a_maxwellian():
    from dgbgk import meshgen
    from oracle_py import OracleRun
    case = _fine_case(meshgen.unit_square(3, left=2, right=2, bottom=2, top=2))
    o = OracleRun(case)
    n, u, v, T = 1.3, 0.2, -0.1, 0.8
    o.set_f(_maxwellian(o, n, u, v, T))
    nf = o.node_fields()
    tol = 1e-5
    assert np.allclose(nf["disnd"], n, rtol=tol)
    assert np.allclose(nf["disux"], u, atol=tol)
    assert np.allclose(nf["disuy"], v, atol=tol)
    assert np.allclose(nf["disps"], n * 1.0 * T, rtol=tol)  # p = n m R T, m = M/NA = 1
    pf = o.point_fields()
    assert np.allclose(pf["nd"], n, rtol=tol)
    assert np.allclose(pf["temp"], T, rtol=tol)
    assert np.allclose(pf["uvel"], u, atol=tol)
    # COMPMU: mu = sqrt(2) * (3.1416 d^2) * n * sqrt(pi)/(2 beta)
    o.prepare_step()
    mu = o.node_fields()["mu"]
    beta = math.sqrt(1.0 / (2 * T))
    d = case.prm["d"]
    expect = math.sqrt(2.0) * (3.1416 * d * d) * n * math.sqrt(math.pi) / (2 * beta)
    assert np.allclose(mu, expect, rtol=1e-4)  # |c-u| has a kink at c = u


def test_point_average_of_a_linear_density():
    """Every DG node on a point carries the same n: the area-weighted point
    average (src/macros.f90:156-185) returns it."""
    from dgbgk import meshgen
    from oracle_py import OracleRun
    m = meshgen.unit_square(6, left=2, right=2, bottom=2, top=2)
    case = _fine_case(m)
    o = OracleRun(case)
    f0 = _maxwellian(o, 1.0, 0.0, 0.0, 1.0)
    nd_pt = 1.0 + 0.5 * m.coord[:, 0] - 0.25 * m.coord[:, 1]
    f = f0 * nd_pt[m.intma - 1][:, None, :]
    o.set_f(f)
    pf = o.point_fields()
    assert np.allclose(pf["nd"], nd_pt, rtol=1e-5)
    assert np.allclose(pf["temp"], 1.0, rtol=1e-5)


def test_flucon2_eta_of_a_wall_maxwellian():
    from dgbgk import meshgen
    from oracle_py import OracleRun
    m = meshgen.unit_square(4, left=2, right=2, bottom=4, top=4, tw=1)
    case = _fine_case(m, rv=8.0)
    o = OracleRun(case)
    n, R, Tw = 0.7, 1.0, 1.0
    o.set_f(_maxwellian(o, n, 0.0, 0.0, Tw))
    o.prepare_step()
    eta = o.eta()
    wall = m.bsido[:, 3] == 4
    beta = math.sqrt(1.0 / (2 * R * Tw))
    flux = n / (2 * math.sqrt(math.pi) * beta)
    expect = flux / math.sqrt(2 * 3.1416 * R ** 3 * Tw ** 3)
    assert np.allclose(eta[wall], expect, rtol=3e-3)
    assert np.all(eta[~wall] == 0.0)


def test_free_stream_preservation_all_outflow():
    """mu = 0 (d = 0), uniform f per ordinate, outflow on every side: the DG
    residual of a constant state vanishes (volume term and upwind edge
    fluxes cancel by the divergence theorem), so f does not change."""
    from dgbgk import cases, meshgen
    from oracle_py import OracleRun
    m = meshgen.body_in_box(4.0, 2.0, 2.0, 1.0, 0.5, h0=0.08, growth=0.2, hmax=0.2, seed=3,
                            wall_type=2, far_types=(2, 2, 2, 2))
    eta, w = cases.midpoint_speeds(6)
    prm = cases.base_params(u0=0.3, kn=1.0, length=1.0, rv=4.0, alpha=1.0)
    prm["d"] = 0.0
    case = cases.Case("fs", m, eta, w, 12, prm)
    o = OracleRun(case)
    f0 = o.get_f()
    o.step(3)
    f1 = o.get_f()
    assert np.max(np.abs(f1 - f0)) <= 1e-13 * np.max(np.abs(f0))


def test_interior_fluxes_cancel_pairwise():
    """With no boundary flux (every side an inflow side, which carries no
    flux) the element RHS after the edge loop sums to the volume-term sum,
    and the volume term alone sums to zero (sum of the P1 gradients is 0)."""
    from dgbgk import cases, meshgen
    from oracle_py import OracleRun
    m = meshgen.body_in_box(3.0, 2.0, 1.5, 1.0, 0.5, h0=0.1, growth=0.2, hmax=0.25, seed=4,
                            wall_type=1, far_types=(1, 1, 1, 1))
    eta, w = cases.midpoint_speeds(4)
    prm = cases.base_params(u0=0.3, kn=0.3, length=1.0, rv=4.0, alpha=1.0)
    case = cases.Case("cancel", m, eta, w, 8, prm)
    o = OracleRun(case)
    f = o.get_f()
    rng = np.random.default_rng(2)
    o.set_f(f * (1 + 0.3 * rng.random(f.shape)))
    dte = o.prepare_step()
    for iv in (1, 7, 20, 32):
        st = o.debug_ordinate(iv, dte)
        scale = np.abs(st["rhs_elc"]).sum()
        assert abs(st["rhs_elc"].sum()) <= 1e-12 * scale
        assert abs(st["rhs_edge"].sum() - st["rhs_elc"].sum()) <= 1e-12 * scale
        assert not np.allclose(st["rhs_edge"], st["rhs_elc"])  # the edges did act


def test_inflow_value_closed_form():
    """INFLOW's Maxwellian (src/inflow.f90:28-39) in the non-dimensional
    deck: rho = p = 1, so f_inf(c) = (1/(2 pi_ref)) exp(-|c-u|^2/2)."""
    from oracle_py import lib
    cinf = np.array([1.0, 1e-5, 0.3, -0.2])
    for ux, uy in ((0.0, 0.0), (1.5, -0.7)):
        val = lib().ora_inflow_value(cinf, 1.0, 6.022e26, ux, uy)
        expect = (0.5 / 3.1416) * math.exp(-0.5 * ((ux - 0.3) ** 2 + (uy + 0.2) ** 2))
        assert abs(val - expect) <= 1e-14 * expect


def test_mass_budget_closes():
    """The reference method conserves nothing exactly (clamp, inflow reset,
    discrete Maxwellian): the per-step total mass is a smooth record whose
    change is small and O(dt)."""
    from small_cases import c1
    from oracle_py import OracleRun
    o = OracleRun(c1())
    r = o.step(5)
    m = np.array([x["mass"] for x in r])
    assert np.all(np.abs(np.diff(m)) < 1e-2 * m[:-1])


@pytest.mark.parametrize("faithful", [0, 1])
def test_search_tables_equal_literal_scans(faithful):
    """The first-match tables that replace the reference's O(N^2) searches
    give the same steps as the literal scans."""
    from small_cases import cylinder_small
    from oracle_py import OracleRun
    case = cylinder_small(elems=600)
    a = OracleRun(case, faithful=0)
    b = OracleRun(case, faithful=faithful)
    assert a.step(3) == b.step(3)
    assert np.array_equal(a.get_f(), b.get_f())
