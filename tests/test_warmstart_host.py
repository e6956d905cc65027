"""Host-side pieces of the learned warm start (no GPU): the case
distribution, the feature and target maps, and the far field."""
import math

import numpy as np

from dgbgk import cases, meshgen, warmstart


def test_draw_covers_the_stated_ranges_and_is_seeded():
    a = warmstart.draw(4000, seed=7)
    b = warmstart.draw(4000, seed=7)
    assert [(s.mach, s.kn) for s in a] == [(s.mach, s.kn) for s in b]
    mach = np.array([s.mach for s in a])
    kn = np.array([s.kn for s in a])
    assert 0.1 <= mach.min() and mach.max() <= 1.5
    assert warmstart.KN_MIN <= kn.min() and kn.max() <= 1.0
    # log-uniform: log10(Kn) has the uniform mean
    lo = math.log10(warmstart.KN_MIN)
    assert abs(np.log10(kn).mean() - 0.5 * lo) < 0.03
    assert abs(mach.mean() - 0.8) < 0.03


def test_targets_and_untarget_are_inverse():
    rng = np.random.default_rng(1)
    s = warmstart.Sample(0.7, 0.2)
    n = 50
    fields = dict(rho=rng.uniform(0.5, 2, n), uvel=rng.normal(size=n),
                  vvel=rng.normal(size=n), temp=rng.uniform(0.8, 1.5, n))
    y = warmstart.targets(fields, s)
    rho, u, v, t = warmstart.untarget(y, s)
    for a, b in ((rho, fields["rho"]), (u, fields["uvel"]), (v, fields["vvel"]),
                 (t, fields["temp"])):
        assert np.allclose(a, b, rtol=2e-7, atol=1e-6)
    # the seed state is clamped positive
    rho, _, _, t = warmstart.untarget(np.full((3, 4), -1.0, np.float32), s)
    assert rho.min() > 0 and t.min() > 0


def test_features_and_far_field():
    mesh = warmstart.mesh_for(600)
    types = set(int(t) for t in mesh.bsido[:, 3])
    assert types == {1, 4}  # freestream box, diffuse cylinder
    wd = meshgen.wall_distance(mesh)
    x = warmstart.features(mesh, wd, warmstart.Sample(0.5, 0.1))
    assert x.shape == (mesh.npoin, 5) and x.dtype == np.float32
    assert 0.0 <= x[:, 0].min() and x[:, 0].max() <= 1.0
    assert 0.0 <= x[:, 1].min() and x[:, 1].max() <= 1.0
    assert np.allclose(x[:, 3], 0.5) and np.allclose(x[:, 4], -1.0)
    # wall distance: zero on the cylinder, 4.5 at the box's bottom-centre
    # (the cylinder of radius 0.5 sits at (5, 5))
    on_wall = mesh.bsido[mesh.bsido[:, 3] == 4][:, :2].ravel() - 1
    assert np.abs(wd[on_wall]).max() < 1e-12
    r = np.hypot(mesh.coord[:, 0] - 5.0, mesh.coord[:, 1] - 5.0) - 0.5
    assert np.abs(wd - r).max() < 0.02  # segments approximate the circle


def test_case_uses_gauss_legendre_speeds_and_flow_parameters():
    mesh = warmstart.mesh_for(600)
    c = warmstart.cylinder_case(mesh, warmstart.Sample(1.0, 0.3), 8, 16)
    eta, w = cases.gauss_legendre_speeds(8)
    assert np.array_equal(c.eta, eta) and np.array_equal(c.w, w)
    assert c.prm["u0"] == cases.mach_to_u(1.0) and c.prm["alpha"] == 1.0
    assert c.prm["rv"] == cases.rv_for(cases.mach_to_u(1.0))
