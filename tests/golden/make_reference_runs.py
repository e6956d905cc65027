"""Records the reference program's own output on small decks (golden runs).

Runs the reference's whole program, built from its unmodified sources by
oracle/build_ref_program.sh (default REAL(4), as the reference is built),
at one MPI rank on two decks whose velocity grid is the reference's own
Lobatto10 file with TORDER = 10 (the grids on which its wall ring lookup,
src/getbou.f90:223-243, is defined):

  lobatto10_walls     unit square 12 x 12 (288 triangles), inflow left,
                      outflow right, diffuse/specular walls (alpha 0.9)
                      top and bottom;
  cylinder_lobatto10  cylinder in a 20 x 10 box (2021 triangles), wall
                      alpha 0.9;
  lobatto10_walls_immat0  the first with IMMAT = 0, the consistent-mass
                      branch (CONMAT, CMMINC).

Each runs 1010 steps with FORCEOUT = 1, so RESIDUAL.RES holds every step's
LOG(SUMRES/SRITM1) with the lift and drag of the last gathered pressure
(the initial state up to step 999, the step-1000 state from then on:
src/advnce.f90:403-435), and RESULTS1/2.RES hold the point fields at steps
0 and 1000 (src/advnce.f90:163-170,405-416).  The decks, meshes and
velocity files are written by this project's own writers (dgbgk.io), and
inputs and outputs are kept under tests/golden/ref_runs/<deck>/ with the
run's wall time (timing.json).

Usage (in a container holding the reference tree):
    bash oracle/build_ref_program.sh
    python tests/golden/make_reference_runs.py
"""
from __future__ import annotations

import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "dg-bgk_amd"))
sys.path.insert(0, os.path.dirname(HERE))

from dgbgk import io  # noqa: E402
import small_cases  # noqa: E402

BGK = os.path.join(ROOT, "oracle", "_ref", "bgk32", "BGK")
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
def _lobatto10_walls_immat0():
    case = small_cases.lobatto10_walls()
    case.prm["immat"] = 0
    return case


DECKS = {"lobatto10_walls": small_cases.lobatto10_walls,
         "cylinder_lobatto10": small_cases.cylinder_lobatto10,
         "lobatto10_walls_immat0": _lobatto10_walls_immat0}
NTIME = 1010
KEEP = ("run.con", "mesh.plt", "velocity.txt", "RESIDUAL.RES", "RESULTS1.RES",
        "RESULTS2.RES", "FILE.OUT")


def cpu_model() -> str:
    for line in open("/proc/cpuinfo"):
        if line.startswith("model name"):
            return line.split(":", 1)[1].strip()
    return platform.processor()


def run(name: str, make) -> dict:
    case = make()
    out = os.path.join(HERE, "ref_runs", name)
    with tempfile.TemporaryDirectory() as tmp:
        deck = io.write_run_dir(case, tmp, ntime=NTIME, forceout=1)
        t0 = time.perf_counter()
        # the reference keeps ~100 MB of fixed-size locals on the stack
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited && {MPIEXEC} -np 1 {BGK} "
                            f"{os.path.basename(deck)}"],
                           cwd=tmp, capture_output=True, text=True, timeout=1800)
        wall = time.perf_counter() - t0
        if r.returncode != 0 or "END OF TIME LOOP" not in r.stdout:
            raise RuntimeError(f"reference run {name} failed:\n{r.stdout[-2000:]}{r.stderr}")
        steps = [float(line.split(":", 1)[1]) for line in r.stdout.splitlines()
                 if "Iteration time:" in line]
        os.makedirs(out, exist_ok=True)
        for f in KEEP:
            shutil.copy(os.path.join(tmp, f), os.path.join(out, f))
    dofv = 3 * case.mesh.nelem * case.nvel
    t = {"deck": name, "nelem": case.mesh.nelem, "npoin": case.mesh.npoin,
         "nboun": case.mesh.nboun, "ordinates": case.nvel, "steps": NTIME,
         "wall_s": wall, "sum_iteration_time_s": sum(steps),
         "dofv_per_s": dofv * NTIME / sum(steps) if sum(steps) else None,
         "mpi_ranks": 1, "cores": 1, "precision": "REAL(4) as built",
         "cpu": cpu_model(),
         "compiler": subprocess.run(["amdflang", "--version"], capture_output=True,
                                    text=True).stdout.splitlines()[0]}
    with open(os.path.join(out, "timing.json"), "w") as fh:
        json.dump(t, fh, indent=1)
    return t


def main():
    if not os.path.exists(BGK):
        raise SystemExit(f"{BGK} not built: run bash oracle/build_ref_program.sh")
    only = sys.argv[1:]
    for name, make in DECKS.items():
        if not only or name in only:
            print(json.dumps(run(name, make)))


if __name__ == "__main__":
    main()
