"""Small decks the oracle finishes in seconds, shared by CPU and GPU tests."""
import os

import numpy as np

from dgbgk import cases, io, meshgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def c1(n=32):
    """Config 1 (unit square, specular walls, 8 x 16 ordinates)."""
    return cases.unit_square_channel(n=n)


def cylinder_small(elems=2000, rorder=8, torder=16):
    """Cylinder flow with diffuse type-4 walls (FLUCON2 and reflections)."""
    return cases.cylinder_flow(target_elems=elems, rorder=rorder, torder=torder)


def lobatto10_walls(n=12, alpha=0.9):
    """Unit square with type-4 walls top and bottom and the reference's own
    Lobatto10 file with TORDER = 10: a grid on which the reference's ring
    lookup is defined (RORDER == TORDER, first node at -1)."""
    eta, w = io.read_lobatto(os.path.join(GOLDEN, "Lobatto10.txt"))
    m = meshgen.unit_square(n, left=1, right=2, bottom=4, top=4, tw=1, alpha100=90)
    prm = cases.base_params(u0=0.4, kn=0.2, length=1.0, rv=5.0, alpha=alpha)
    return cases.Case("lobatto10_walls", m, eta, w, 10, prm, nsteps=10)


def cylinder_lobatto10(elems=1500, alpha=0.9):
    eta, w = io.read_lobatto(os.path.join(GOLDEN, "Lobatto10.txt"))
    c = cases.cylinder_flow(target_elems=elems, rorder=10, torder=10)
    prm = cases.base_params(u0=0.6, kn=0.1, length=1.0, rv=5.5, alpha=alpha)
    return cases.Case("cylinder_lobatto10", c.mesh, eta, w, 10, prm, nsteps=10)


def shuffled(case, seed=0):
    """case on its mesh with the elements in a random order; returns the case
    and perm (element i of the new mesh is element perm[i] of case's)."""
    m, perm = meshgen.shuffle_elements(case.mesh, seed)
    return cases.Case(case.name + "-shuffled", m, case.eta, case.w, case.torder,
                      dict(case.prm), nsteps=case.nsteps), perm


def rel_l2(a, b):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_l2_vec(ax, ay, bx, by):
    """Relative L2 of a 2-vector field (components compared against the
    norm of the whole vector, so a near-zero component is not divided by
    its own rounding noise)."""
    ax, ay, bx, by = (np.asarray(x, float).ravel() for x in (ax, ay, bx, by))
    num = np.sqrt(np.sum((ax - bx) ** 2 + (ay - by) ** 2))
    den = np.sqrt(np.sum(bx ** 2 + by ** 2))
    return float(num / max(den, 1e-300))
