"""The reference's file formats: run deck (.con), mesh (.plt), radial
quadrature (Lobatto file) and the output files RESIDUAL.RES, RESULTS1.RES,
RESULTS2.RES and FILE.OUT.

Readers accept what the reference's list-directed READs accept for these
files; writers produce the reference's layouts (src/advnce.f90:429-432,
488-489; src/spectralbgk.f90:240-282).
"""
from __future__ import annotations

import os
import re

import numpy as np

from .meshgen import Mesh

# Defaults of the /Input/ namelist (src/spectralbgk.f90:116-147).
DECK_DEFAULTS = {
    "TORDER": 20, "NTIME": 10000, "FORCEOUT": 0, "IMMAT": 1, "INF": 1, "RS": 0,
    "CSAFM": 0.5, "RV": 2000.0, "T1": 293.0, "P1": 0.1, "U0": 100.0, "V0": 0.0,
    "W": 1.0, "ALPHA": 0.9, "R": 287.0, "D": 250e-12, "M": 32.0,
    "CINF": [294.0, 0.001473, 171.85, 0.0], "NVSPACEPART": 1,
    "LOBATTOFILE": "Lobatto20.txt", "PSPACEFILE": "AEROFOIL.RES", "OUTFILE": "FILE.OUT",
    "PARTITIONFILE": "METIS3WAYPARTAEROFOIL13.RES", "RESTARTINFILE": "RESTART.RES",
    "RESIDUALFILE": "RESIDUAL.RES", "RESULTSFILE1": "RESULTS1.RES",
    "RESULTSFILE2": "RESULTS2.RES", "RESTARTOUTFILE": "RESTART.RES",
    "GIDMESHFILE": "GIDMESH.RES",
}


def _fnum(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def read_deck(path: str) -> dict:
    """Parses the &Input ... / namelist of IVD%KEY = value entries."""
    deck = {k: (list(v) if isinstance(v, list) else v) for k, v in DECK_DEFAULTS.items()}
    text = open(path).read()
    m = re.search(r"&\s*input(.*?)(^\s*/|/\s*$)", text, re.I | re.S | re.M)
    body = m.group(1) if m else text
    body = "\n".join(line.split("!")[0] for line in body.splitlines())
    for mm in re.finditer(r"IVD%\s*([A-Za-z0-9_]+)\s*(\(\s*(\d+)\s*\))?\s*=\s*"
                          r"(\"[^\"]*\"|'[^']*'|[^,\n]+)", body, re.I):
        key, idx, val = mm.group(1).upper(), mm.group(3), mm.group(4).strip()
        if val[:1] in "\"'":
            v = val[1:-1]
        else:
            v = val.rstrip(",").strip()
            try:
                v = _fnum(v)
            except ValueError:
                pass
        if idx is not None:
            deck[key][int(idx) - 1] = float(v)
        elif key in deck and isinstance(deck[key], int) and isinstance(v, float):
            deck[key] = int(v)
        else:
            deck[key] = v
    return deck


def read_lobatto(path: str):
    """Line 1: RORDER; then RORDER lines 'eta weight' (src/vspace.f90:16-18)."""
    toks = open(path).read().split()
    r = int(float(toks[0]))
    vals = np.array([_fnum(t) for t in toks[1:1 + 2 * r]], dtype=np.float64).reshape(r, 2)
    return vals[:, 0].copy(), vals[:, 1].copy()


def write_lobatto(path: str, eta, w):
    with open(path, "w") as fh:
        fh.write(f"{len(eta)}\n")
        for a, b in zip(eta, w):
            fh.write(f"{float(a)!r}\t{float(b)!r}\n")


def read_plt(path: str) -> Mesh:
    """Mesh file read as in src/spectralbgk.f90:240-272 and
    src/gtinpt.f90:24-50: NLINES, NLINES title lines, a text line, NDIMN,
    a text line, 'NELEM NPOIN NBOUN', then the connectivity, coordinate and
    boundary-side blocks, each after one header line."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    it = iter(lines)

    def nxt():
        return next(it)

    nlines = int(nxt().split()[0])
    for _ in range(nlines):
        nxt()
    nxt()
    ndimn = int(nxt().split()[0])
    if ndimn != 2:
        raise ValueError("only 2-D meshes")
    nxt()
    ne, npn, nb = (int(x) for x in nxt().split()[:3])
    nxt()
    intma = np.empty((ne, 3), np.int32)
    for i in range(ne):
        t = nxt().split()
        intma[i] = [int(t[1]), int(t[2]), int(t[3])]
    nxt()
    coord = np.empty((npn, 2))
    for i in range(npn):
        t = nxt().split()
        coord[i] = [_fnum(t[1]), _fnum(t[2])]
    nxt()
    bsido = np.empty((nb, 6), np.int32)
    for i in range(nb):
        t = nxt().split()
        bsido[i] = [int(float(x)) for x in t[:6]]
    return Mesh(intma=intma, coord=coord, bsido=bsido, name=os.path.basename(path))


def write_plt(path: str, mesh: Mesh, title: str = "title"):
    with open(path, "w") as fh:
        fh.write("           1\n")
        fh.write(f"{title}\n")
        fh.write("number of dimensions\n           2\n")
        fh.write("nelem   npoin   nboun\n")
        fh.write(f"{mesh.nelem:12d}{mesh.npoin:12d}{mesh.nboun:12d}\n")
        fh.write("  connectivities \n")
        for i, t in enumerate(mesh.intma):
            fh.write(f"{i + 1:10d}{t[0]:10d}{t[1]:10d}{t[2]:10d}         0\n")
        fh.write("  coordinates \n")
        for i, c in enumerate(mesh.coord):
            fh.write(f"{i + 1:10d} {float(c[0])!r} {float(c[1])!r}\n")
        fh.write("  boundary sides \n")
        for b in mesh.bsido:
            fh.write(" ".join(f"{int(x):10d}" for x in b) + "\n")


def write_deck(path: str, keys: dict):
    """Writes an &Input namelist deck (quoted strings, as gfortran/flang need)."""
    with open(path, "w") as fh:
        fh.write("&Input\n")
        for k, v in keys.items():
            if isinstance(v, (list, tuple, np.ndarray)):
                for i, x in enumerate(v):
                    fh.write(f"IVD%{k}({i + 1}) = {float(x)!r},\n")
            elif isinstance(v, str):
                fh.write(f'IVD%{k} = "{v}",\n')
            elif isinstance(v, (int, np.integer)):
                fh.write(f"IVD%{k} = {int(v)},\n")
            else:
                fh.write(f"IVD%{k} = {float(v)!r},\n")
        fh.write("/\n")


def fortran_e(x: float) -> str:
    """Fortran 1PE12.5 edit descriptor."""
    if x == 0.0:
        return "  0.00000E+00"
    s = f"{x:.5E}"
    mant, exp = s.split("E")
    e = int(exp)
    return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}".rjust(12)


def results_block(itime: int, a, b, c) -> str:
    """One RESULTS1/RESULTS2 block (src/advnce.f90:488-489)."""
    out = [f"RESULTS        1   {itime:5d}    2      1      0"]
    for ip in range(len(a)):
        out.append(f"{ip + 1:5d}    {fortran_e(a[ip])} {fortran_e(b[ip])} {fortran_e(c[ip])} ")
    return "\n".join(out) + "\n"


def read_residual(path: str) -> np.ndarray:
    """RESIDUAL.RES: one row per step, ITIME LOG(SUMRES/SRITM1) [LIFT DRAG]
    (src/advnce.f90:427-432, list-directed)."""
    rows = [[_fnum(t) for t in line.split()] for line in open(path) if line.strip()]
    return np.array(rows, dtype=np.float64)


def read_results(path: str) -> dict:
    """RESULTS1.RES / RESULTS2.RES: {ITIME: (NPOIN, 3) array} of the blocks
    written at output steps (src/advnce.f90:163-170,405-416,466-470)."""
    blocks, cur = {}, None
    for line in open(path):
        if line.startswith("RESULTS"):
            cur = int(line.split()[2])
            blocks[cur] = []
        elif line.strip():
            blocks[cur].append([_fnum(t) for t in line.split()[1:4]])
    return {k: np.array(v, dtype=np.float64) for k, v in blocks.items()}


def _ordinate_xy(eta, torder: int, rv: float):
    """Cartesian velocity of every ordinate, ring-major, as the reference's
    restart writer maps VCORD back (src/output.f90:112-118, PI = 3.1416)."""
    eta = np.asarray(eta, float)
    iz = np.arange(1, torder + 1)
    rt = np.repeat(eta * (rv / 2) + rv / 2, torder)
    th = np.tile(-1.0 + (2.0 * iz) / torder, len(eta)) * 3.1416
    return rt * np.cos(th), rt * np.sin(th)


def write_restart(path: str, f, npoin: int, eta, torder: int, rv: float):
    """f (E, V, 3) in the reference writer's layout (src/output.f90:54-58,
    110-125): 'NPOIN     VNPNT', the two counts, then for every ordinate
    and element 'IE CX CY f1 f2 f3', with f at full double precision."""
    f = np.asarray(f, float)
    E, V, _ = f.shape
    cx, cy = _ordinate_xy(eta, torder, rv)
    with open(path, "w") as fh:
        fh.write("NPOIN     VNPNT\n")
        fh.write(f" {npoin}     {V}\n")
        ie = np.arange(1, E + 1)
        for v in range(V):
            rows = np.column_stack([ie, np.full(E, cx[v]), np.full(E, cy[v]), f[:, v, :]])
            np.savetxt(fh, rows, fmt=["%10d", "%15.7E", "%15.7E", "%25.17E", "%25.17E",
                                      "%25.17E"], delimiter=" ")


def read_restart(path: str, nelem: int, nvel: int, npoin: int) -> np.ndarray:
    """f (E, V, 3) from a restart file of write_restart's layout: the last
    three numbers of each line are the element's three nodal values."""
    with open(path) as fh:
        fh.readline()
        counts = [int(_fnum(t)) for t in fh.readline().split()]
        if counts != [npoin, nvel]:
            raise ValueError(f"restart file is for NPOIN, VNPNT = {counts}, "
                             f"not {[npoin, nvel]}")
        data = np.loadtxt(fh, converters=None, dtype=float)
    data = data.reshape(nvel, nelem, -1)
    return np.ascontiguousarray(data[:, :, -3:].transpose(1, 0, 2))


def write_gid_mesh(path: str, mesh: Mesh):
    """The reference's GiD mesh file (src/output.f90:130-153, formats
    200-240)."""
    with open(path, "w") as fh:
        fh.write("mesh dimension = 2 elemtype triangle nnode = 3\n")
        fh.write("   COORDINATES\n")
        for i, (x, y) in enumerate(mesh.coord, 1):
            fh.write(f"{i:10d}  {x:10.5f}  {y:10.5f}\n")
        fh.write("end coordinates\n")
        fh.write("elements\n")
        for i, t in enumerate(mesh.intma, 1):
            fh.write(f"{i:8d}{t[0]:8d}{t[1]:8d}{t[2]:8d}\n")
        fh.write("end elements\n")


def read_gid_mesh(path: str):
    """(coord (P, 2), intma (E, 3)) from a GiD mesh file."""
    lines = open(path).read().splitlines()
    a = lines.index("   COORDINATES") + 1
    b = lines.index("end coordinates")
    c = lines.index("elements") + 1
    d = lines.index("end elements")
    coord = np.array([[float(t) for t in l.split()[1:3]] for l in lines[a:b]])
    intma = np.array([[int(t) for t in l.split()[1:4]] for l in lines[c:d]], dtype=np.int32)
    return coord, intma


def write_run_dir(case, directory: str, ntime: int, forceout: int = 1,
                  deck_name: str = "run.con", restart_in: str | None = None) -> str:
    """Writes deck + mesh + velocity file of a Case as a reference-style run
    directory; returns the deck path.  restart_in: start from that restart
    file (IVD%RS = 1) instead of the INICON2 state.  The nodal partition
    file names one partition for every point (all zeros), as a one-rank run
    of the reference reads it (src/ptmesh.f90:64-79)."""
    os.makedirs(directory, exist_ok=True)
    write_plt(os.path.join(directory, "mesh.plt"), case.mesh)
    write_lobatto(os.path.join(directory, "velocity.txt"), case.eta, case.w)
    with open(os.path.join(directory, "partition.txt"), "w") as fh:
        fh.write("0\n" * case.mesh.npoin)
    p = case.prm
    keys = {
        "TORDER": int(case.torder), "rv": p["rv"], "FORCEOUT": int(forceout),
        "NTIME": int(ntime), "IMMAT": int(p.get("immat", 1)), "ALPHA": p["alpha"],
        "CSAFM": p["csafm"],
        "RS": 1 if restart_in else 0, "T1": p["t1"], "P1": p["p1"], "U0": p["u0"], "V0": p["v0"],
        "INF": int(p.get("inf", 1)), "CINF": list(p["cinf"]), "NVSPACEPART": 1,
        "R": p["rgas"], "d": p["d"], "M": p["m"],
        "LobattoFile": "velocity.txt", "PSpaceFile": "mesh.plt", "OutFile": "FILE.OUT",
        "PartitionFile": "partition.txt",
        "ResidualFile": "RESIDUAL.RES", "ResultsFile1": "RESULTS1.RES",
        "ResultsFile2": "RESULTS2.RES", "RestartOutFile": "RESTART_OUT.RES",
        "GIDMeshFile": "GIDMESH.RES",
    }
    if restart_in:
        keys["RestartInFile"] = restart_in
    path = os.path.join(directory, deck_name)
    write_deck(path, keys)
    return path
