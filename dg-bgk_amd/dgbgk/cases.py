"""Workloads as reference run decks: mesh + radial quadrature + deck keys.

Flow conditions are non-dimensionalised through the reference's own deck
keys (src/spectralbgk.f90:116-147): gas constant R = 1 and molar mass
M = 6.022e26 make the molecular mass M/NA exactly 1, so number density is
mass density; T = 1 and P = 1e-5 bar (P enters multiplied by 10**5) give
rho = p = 1.  The thermal speed sqrt(RT) is then 1, the 2-D monatomic sound
speed is sqrt(gamma R T) with gamma = 2, and Kn enters through the
hard-sphere diameter d = (sqrt(2) * 3.1416 * n * Kn * L)**-0.5 that makes the
reference's collision frequency (src/compmu.f90:45,90-92) give a mean
free path of Kn * L.

Radial quadrature.  The reference normalises every velocity sum by
FRAC = 4/SUMWEIGHT (src/macros.f90:55), which is right only for weights
summing to 2 (its own Gauss-Lobatto files).  End-point trapezoid weights on
(0, rv] sum to 2 - 1/R and would inflate every moment by 2R/(2R-1), so
config 1's "8 speeds uniform on (0,4]" uses the midpoint rule
eta_i = (2i-1)/R - 1, w_i = 2/R.  The method is not mass-conservative: a
discrete Maxwellian's density differs from its parameter by a quadrature
error eps, and BGK relaxation then grows the density at rate mu*eps.  For
the steady-state workloads (cylinder, particle, learned warm start) eps
must stay well below the outflow rate, so they use Gauss-Legendre radial
nodes (eps ~ 1e-6; the midpoint rule gives 0.1-1%).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import meshgen

GAMMA_2D = 2.0
PI_REF = 3.1416


@dataclass
class Case:
    name: str
    mesh: meshgen.Mesh
    eta: np.ndarray          # radial nodes in [-1, 1]
    w: np.ndarray            # radial weights
    torder: int
    prm: dict = field(default_factory=dict)
    nsteps: int = 200

    @property
    def rorder(self) -> int:
        return int(len(self.eta))

    @property
    def nvel(self) -> int:
        return self.rorder * self.torder

    @property
    def dofv_per_step(self) -> int:
        return 3 * self.mesh.nelem * self.nvel


def midpoint_speeds(r: int):
    i = np.arange(1, r + 1, dtype=np.float64)
    return (2 * i - 1) / r - 1.0, np.full(r, 2.0 / r)


def gauss_legendre_speeds(r: int):
    eta, w = np.polynomial.legendre.leggauss(r)
    return eta.astype(np.float64), w.astype(np.float64)


def diameter_for_kn(kn: float, length: float = 1.0, n: float = 1.0) -> float:
    return (math.sqrt(2.0) * PI_REF * n * kn * length) ** -0.5


def base_params(u0: float, kn: float, length: float, rv: float, alpha: float,
                csafm: float = 0.5, v0: float = 0.0) -> dict:
    return dict(rv=rv, csafm=csafm, alpha=alpha, rgas=1.0, d=diameter_for_kn(kn, length),
                m=6.022e26, t1=1.0, p1=1e-5, u0=u0, v0=v0,
                cinf=[1.0, 1e-5, u0, v0], inf=1)


def mach_to_u(mach: float) -> float:
    return mach * math.sqrt(GAMMA_2D)


def rv_for(u0: float) -> float:
    """Velocity-grid radius: 5 thermal speeds beyond the flow speed."""
    return math.ceil(2.0 * (5.0 + abs(u0))) / 2.0


def unit_square_channel(n: int = 32, rorder: int = 8, torder: int = 16) -> Case:
    """Unit square, 2n^2 P1 elements; 16 angles x 8 speeds on (0,4];
    rho=1, T=1, u=(0.1,0); inflow left, outflow right, specular walls top
    and bottom (type 3, alpha=0); Kn=0.1; 200 steps at CFL 0.5."""
    eta, w = midpoint_speeds(rorder)
    m = meshgen.unit_square(n, left=1, right=2, bottom=3, top=3, tw=1, alpha100=0)
    prm = base_params(u0=0.1, kn=0.1, length=1.0, rv=4.0, alpha=0.0)
    return Case("unit_square_channel", m, eta, w, torder, prm, nsteps=200)


def cylinder_mesh(target_elems: int = 100_000, seed: int = 1,
                  far_types=(1, 2, 2, 2)) -> meshgen.Mesh:
    """20D x 10D box, cylinder of diameter D=1 at (5,5), graded Delaunay.
    far_types: side types of the (left, right, bottom, top) box sides."""
    hmax = 0.068 * math.sqrt(100_000 / target_elems)
    h0 = min(0.02 * math.sqrt(100_000 / target_elems), 0.5 * hmax)
    return meshgen.body_in_box(20.0, 10.0, 5.0, 5.0, 1.0, h0=h0, growth=0.12,
                               hmax=hmax, seed=seed, wall_type=4, tw=1,
                               alpha100=100, far_types=tuple(far_types),
                               name=f"cylinder_{target_elems}")


def cylinder_flow(target_elems: int = 100_000, rorder: int = 16, torder: int = 32,
            mach: float = 0.5, kn: float = 0.05) -> Case:
    """Cylinder flow: Mach 0.5, Kn 0.05 (on D), diffuse isothermal wall T=1,
    512 ordinates, 2000 steps."""
    m = cylinder_mesh(target_elems)
    u0 = mach_to_u(mach)
    eta, w = gauss_legendre_speeds(rorder)
    prm = base_params(u0=u0, kn=kn, length=1.0, rv=rv_for(u0), alpha=1.0)
    return Case("cylinder_flow", m, eta, w, torder, prm, nsteps=2000)


def particle_mesh(target_elems: int = 1_000_000, far_types=(1, 2, 2, 2),
                  h0: float | None = None) -> meshgen.Mesh:
    """200D x 100D box, particle of diameter D = 1 at (60, 50), graded
    Delaunay; h0 the spacing on the particle (default 0.02, or hmax/4 on
    coarse meshes).  far_types: (left, right, bottom, top) box side types."""
    hmax = 0.215 * math.sqrt(1_000_000 / target_elems)
    return meshgen.body_in_box(200.0, 100.0, 60.0, 50.0, 1.0,
                               h0=min(0.02, hmax / 4) if h0 is None else h0,
                               growth=0.1, hmax=hmax, seed=5, wall_type=4, tw=1,
                               alpha100=100, far_types=tuple(far_types),
                               name=f"particle_{target_elems}")


def particle_drag(target_elems: int = 1_000_000, rorder: int = 32, torder: int = 64,
            mach: float = 0.3, kn: float = 1.0, far_types=(1, 2, 2, 2),
            h0: float | None = None) -> Case:
    """Particle of diameter D in a 200D x 100D box, ~1M graded triangles,
    2048 ordinates, Mach 0.3, Kn 1."""
    m = particle_mesh(target_elems, far_types, h0)
    u0 = mach_to_u(mach)
    eta, w = gauss_legendre_speeds(rorder)
    prm = base_params(u0=u0, kn=kn, length=1.0, rv=rv_for(u0), alpha=1.0)
    return Case("particle_drag", m, eta, w, torder, prm, nsteps=100)


def lobatto_case(mesh: meshgen.Mesh, eta, w, torder: int, alpha: float = 0.9,
                 u0: float = 0.3, kn: float = 0.1, rv: float = 5.0) -> Case:
    """Any mesh with a given radial quadrature (e.g. a Lobatto file)."""
    prm = base_params(u0=u0, kn=kn, length=1.0, rv=rv, alpha=alpha)
    return Case("custom", mesh, np.asarray(eta, float), np.asarray(w, float),
                torder, prm)


def from_run_dir(directory: str, deck: str = "run.con", single: bool = False) -> Case:
    """The Case a reference run directory describes: its deck's keys, its
    mesh file and its velocity file.  single=True rounds every real input
    to REAL(4), as the reference (default REAL) reads them."""
    import os

    from . import io
    keys = io.read_deck(os.path.join(directory, deck))
    m = io.read_plt(os.path.join(directory, keys["PSPACEFILE"]))
    eta, w = io.read_lobatto(os.path.join(directory, keys["LOBATTOFILE"]))
    r = (lambda x: float(np.float32(x))) if single else float
    if single:
        m.coord = m.coord.astype(np.float32).astype(np.float64)
        eta = eta.astype(np.float32).astype(np.float64)
        w = w.astype(np.float32).astype(np.float64)
    prm = dict(rv=r(keys["RV"]), csafm=r(keys["CSAFM"]), alpha=r(keys["ALPHA"]),
               rgas=r(keys["R"]), d=r(keys["D"]), m=r(keys["M"]), t1=r(keys["T1"]),
               p1=r(keys["P1"]), u0=r(keys["U0"]), v0=r(keys["V0"]),
               cinf=[r(x) for x in keys["CINF"]], inf=int(keys["INF"]),
               immat=int(keys.get("IMMAT", 1)))
    return Case(os.path.basename(os.path.normpath(directory)), m, eta, w,
                int(keys["TORDER"]), prm, nsteps=int(keys["NTIME"]))
