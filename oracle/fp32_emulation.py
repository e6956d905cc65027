"""Precision study of the FP32-bulk form with the oracle's FP32 emulation.

TEST INFRASTRUCTURE ONLY (a study script; nothing in the product uses it).

Runs the oracle on one deck with chosen stages of the step emulated in
FP32 (ora_set_emulation: f stored rounded to nearest or with the +-1/2 ulp
dither, the Maxwellian parameters and exponential, the flux/predictor/
increment arithmetic, the time step) and reports the relative L2 error of
the point fields against the plain FP64 oracle after N steps, one process
per configuration.  It separates what storing f in FP32 costs from what
FP32 arithmetic adds.

    python oracle/fp32_emulation.py [--steps 2000] [--deck c1gl|c1|cylinder]
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "dg-bgk_amd"))

import oracle_py as op  # noqa: E402

CONFIGS = {
    "store_rn": op.EMU_STORE_RN,
    "store_dither": op.EMU_STORE_SR,
    "params": op.EMU_PARAMS,
    "arith": op.EMU_ARITH | op.EMU_DTE,
    "arith_predictor": op.EMU_PRED,
    "arith_edges": op.EMU_EDGE,
    "arith_increment": op.EMU_INC,
    "params+arith": op.EMU_PARAMS | op.EMU_ARITH | op.EMU_DTE,
    "arith+form": op.EMU_ARITH | op.EMU_DTE | op.EMU_FORM,
    "store_dither+params+arith+form": (op.EMU_STORE_SR | op.EMU_PARAMS | op.EMU_ARITH |
                                       op.EMU_DTE | op.EMU_FORM),
    "store_dither+params+increment": op.EMU_STORE_SR | op.EMU_PARAMS | op.EMU_INC | op.EMU_DTE,
    "store_rn+params+increment": op.EMU_STORE_RN | op.EMU_PARAMS | op.EMU_INC | op.EMU_DTE,
    "store_rn+params+arith": op.EMU_STORE_RN | op.EMU_PARAMS | op.EMU_ARITH | op.EMU_DTE,
    "store_dither+params+arith": op.EMU_STORE_SR | op.EMU_PARAMS | op.EMU_ARITH | op.EMU_DTE,
}


def deck(name: str):
    from dgbgk import cases
    if name == "c1":
        return cases.unit_square_channel(n=32)
    if name == "cylinder":
        return cases.cylinder_flow(target_elems=2000, rorder=8, torder=16)
    base = cases.unit_square_channel(n=32)        # the unit square, Gauss-Legendre speeds
    eta, w = cases.gauss_legendre_speeds(len(base.eta))
    return cases.Case("c1_gauss_legendre", base.mesh, eta, w, base.torder, base.prm)


def fields(o):
    p = o.point_fields()
    return {k: p[k].copy() for k in ("rho", "uvel", "vvel", "ps", "temp")}


def run(args):
    name, flags, deckname, steps, checkpoints = args
    case = deck(deckname)
    o = op.OracleRun(case, lookup=1)
    if flags:
        o.set_emulation(flags)
    else:
        o.set_threads(4)
    out, done, masses = {}, 0, []
    for n in checkpoints:
        masses += [r["mass"] for r in o.step(n - done)]
        done = n
        out[n] = fields(o)
    o.close()
    return name, out, np.array(masses)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--deck", default="c1gl", choices=["c1gl", "c1", "cylinder"])
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--only", default="", help="comma-separated configuration names")
    args = ap.parse_args()
    cps = sorted({min(200, args.steps), args.steps})
    names = [n for n in CONFIGS if not args.only or n in args.only.split(",")]
    jobs = [("fp64", 0, args.deck, args.steps, cps)] + [
        (n, CONFIGS[n], args.deck, args.steps, cps) for n in names]
    with mp.Pool(min(args.procs, len(jobs))) as pool:
        res = {n: (f, m) for n, f, m in pool.map(run, jobs)}
    ref, mref = res.pop("fp64")
    table = {}
    for n, (f, m) in res.items():
        row = {}
        for cp in cps:
            a, b = f[cp], ref[cp]
            uv = float(np.sqrt(np.sum((a["uvel"] - b["uvel"]) ** 2 + (a["vvel"] - b["vvel"]) ** 2))
                       / np.sqrt(np.sum(b["uvel"] ** 2 + b["vvel"] ** 2)))
            row[cp] = {"rho": rel(a["rho"], b["rho"]), "temp": rel(a["temp"], b["temp"]),
                       "ps": rel(a["ps"], b["ps"]), "uv": uv}
        row["mass_max_rel"] = float(np.max(np.abs(m - mref) / np.abs(mref)))
        table[n] = row
        print(n, json.dumps(row), flush=True)
    return table


if __name__ == "__main__":
    main()
