"""Learned warm start for steady cylinder flows.

1. Steady states of the cylinder geometry (20D x 10D box, cylinder D at
   (5D, 5D), freestream held on all four box sides) for (Mach, Kn) drawn
   uniform in [0.1, 1.5] and log-uniform in [0.05, 1], each run by this
   solver from the freestream until it is steady.
2. A coordinate MLP (x, y, wall distance, Mach, Kn) -> (rho, u, v, T),
   trained on the device with the BF16 MFMA kernels (dgbgk.mlp).
3. Held-out cases run from the freestream (the method's own start) and from
   the Maxwellian of the network's prediction (dgbgk_seed_maxwellian), in
   FP64 and in the FP32-bulk form.  The unmodified solver judges every run.

Steady means: the step change max|f^{k+1} - f^k| has fallen below tol
times the step-1 change of the FP64 freestream run of the same case.  The
reference's own residual (SUMRES, src/advnce.f90:421-429) cannot serve:
it counts the increment at nodes that INFLOW then resets to the inflow
state (src/inflow.f90:43-59), so it levels off, and it is normalised by
the run's own first step, which a warm start would shrink.

The box sides are inflow (type 1) rather than outflow (type 2): the
reference's outflow side takes the element's own values as the upwind
state also for ordinates entering the domain, and in subsonic flow that
feeds a growing mode along the side, in this solver and in the FP64 CPU
restatement alike.  Kn starts at 0.05: below it, at Mach >= 0.8, the
collision frequency sets the time step at the reference's limit
MU * DTE = 0.9 (src/alotim2.f90:10-15) and the explicit step diverges
within a few hundred steps, again in both.

  python -m dgbgk.warmstart generate --n 256 --seed 1 --data train.npz
  python -m dgbgk.warmstart train --data train.npz --params params.npz
  python -m dgbgk.warmstart evaluate --params params.npz --n 32 --seed 2 --out res.json
"""
from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _lib, cases, meshgen

FEATURES = ("x", "y", "wall_distance", "mach", "kn")
# network outputs: the departures from the freestream (rho = T = 1 in these
# units), scaled by Mach^2 for rho and T (they vanish like Mach^2), so that
# low-Mach states are learned to the same relative accuracy as high-Mach ones
TARGETS = ("(rho-1)/M^2", "u/U0", "v/U0", "(T-1)/M^2")
FAR_FIELD = (1, 1, 1, 1)
KN_MIN = 0.05


@dataclass
class Sample:
    mach: float
    kn: float


@dataclass(frozen=True)
class Geometry:
    """A body-in-box geometry of the warm start: the box (for the feature
    scaling), the wall-distance scale, the spacing on the body, and the
    (Mach, Kn) ranges its training states are drawn from."""
    box: tuple
    wd_scale: float
    mach: tuple
    kn: tuple
    h0: float | None = None


GEOMETRIES = {
    # the cylinder flow's box: 20D x 10D, cylinder at (5D, 5D)
    "cylinder": Geometry(box=(20.0, 10.0), wd_scale=10.0, mach=(0.1, 1.5), kn=(KN_MIN, 1.0)),
    # the particle-drag box: 200D x 100D, particle at (60D, 50D); Mach and
    # Kn about the large case's (0.3, 1)
    "particle": Geometry(box=(200.0, 100.0), wd_scale=100.0, mach=(0.1, 0.8), kn=(0.2, 2.0),
                         h0=0.1),
}


def draw(n: int, seed: int, geometry: str = "cylinder") -> list[Sample]:
    g = GEOMETRIES[geometry]
    rng = np.random.default_rng(seed)
    mach = rng.uniform(g.mach[0], g.mach[1], n)
    kn = 10 ** rng.uniform(math.log10(g.kn[0]), math.log10(g.kn[1]), n)
    return [Sample(float(m), float(k)) for m, k in zip(mach, kn)]


def mesh_for(elems: int, geometry: str = "cylinder", h0: float | None = None) -> meshgen.Mesh:
    """The geometry's mesh of about `elems` triangles, freestream held on all
    four box sides."""
    if geometry == "cylinder":
        return cases.cylinder_mesh(elems, far_types=FAR_FIELD)
    g = GEOMETRIES[geometry]
    return cases.particle_mesh(elems, far_types=FAR_FIELD, h0=g.h0 if h0 is None else h0)


def cylinder_case(mesh, s: Sample, rorder: int, torder: int) -> cases.Case:
    u0 = cases.mach_to_u(s.mach)
    eta, w = cases.gauss_legendre_speeds(rorder)
    prm = cases.base_params(u0=u0, kn=s.kn, length=1.0, rv=cases.rv_for(u0), alpha=1.0)
    return cases.Case(f"cyl_M{s.mach:.3f}_Kn{s.kn:.4f}", mesh, eta, w, torder, prm)


def features(mesh, wd, s: Sample, geometry: str = "cylinder") -> np.ndarray:
    g = GEOMETRIES[geometry]
    n = mesh.npoin
    return np.stack([mesh.coord[:, 0] / g.box[0], mesh.coord[:, 1] / g.box[1],
                     np.log1p(wd) / math.log1p(g.wd_scale), np.full(n, s.mach),
                     np.full(n, math.log10(s.kn))], axis=1).astype(np.float32)


def fields_scaled(fields: dict, s: Sample) -> np.ndarray:
    """What the training data stores per point: rho, u/U0, v/U0, T."""
    u0 = cases.mach_to_u(s.mach)
    return np.stack([fields["rho"], fields["uvel"] / u0, fields["vvel"] / u0,
                     fields["temp"]], axis=1).astype(np.float32)


def net_targets(y: np.ndarray, s: Sample) -> np.ndarray:
    """Network outputs (TARGETS) from stored rho, u/U0, v/U0, T."""
    m2 = s.mach * s.mach
    y = np.asarray(y, np.float64)
    return np.stack([(y[:, 0] - 1.0) / m2, y[:, 1], y[:, 2], (y[:, 3] - 1.0) / m2],
                    axis=1).astype(np.float32)


def targets(fields: dict, s: Sample) -> np.ndarray:
    """Network outputs (TARGETS) from solver point fields."""
    return net_targets(fields_scaled(fields, s), s)


def untarget(y: np.ndarray, s: Sample):
    """Point-wise rho, u, v, T from network outputs (rho and T clamped
    positive)."""
    u0 = cases.mach_to_u(s.mach)
    m2 = s.mach * s.mach
    y = np.asarray(y, np.float64)
    rho = np.maximum(1.0 + m2 * y[:, 0], 1e-3)
    temp = np.maximum(1.0 + m2 * y[:, 3], 1e-3)
    return rho, y[:, 1] * u0, y[:, 2] * u0, temp


def run_to_steady(case, tol: float, max_steps: int, precision=_lib.DGBGK_FP64, seed=None,
                  norm: float | None = None, chunk: int = 1000, log=None,
                  log_every: int = 20000, ref_tol: float | None = None):
    """Steps until change / norm < tol (norm: this run's step-1 change unless
    given).  Wall time covers the steps up to convergence, scaled from the
    timed chunks.  Returns steps, wall seconds, the change history, point
    fields and forces.  With ref_tol the run goes on past convergence to
    change / norm < ref_tol and also returns that state's drag (ref_drag,
    ref_steps); the fields and forces stay those at tol."""
    from .solver import Solver
    s = Solver(case, precision=precision)
    if seed is not None:
        s.seed(*seed)
    s.sync()
    hist = []
    first, conv, steps, wall, bad = None, None, 0, 0.0, False
    at_tol = None  # (fields, lift, drag) when change first fell below tol
    ref_conv = None
    while steps < max_steps and (conv is None or (ref_tol is not None and ref_conv is None)):
        n = min(chunk, max_steps - steps)
        t0 = time.perf_counter()
        recs = s.step(n)
        wall += time.perf_counter() - t0
        bad = False
        for i, r in enumerate(recs):
            if not math.isfinite(r["mass"]):  # NaN in f (the maxima skip NaN)
                bad = True
                break
            if first is None:
                first = r["change"]
            hist.append(r["change"])
            if conv is None and r["change"] < tol * (norm if norm is not None else first):
                conv = steps + i + 1
            if (ref_tol is not None and ref_conv is None and
                    r["change"] < ref_tol * (norm if norm is not None else first)):
                ref_conv = steps + i + 1
        steps += n
        if bad:
            conv = None
            break
        if conv is not None and at_tol is None:
            at_tol = (s.point_fields(), *s.forces())
            wall_at_tol = wall
        if log is not None and steps % log_every == 0:
            log(f"    ... {steps} steps, change {hist[-1]:.3e} "
                f"(target {tol * (norm if norm is not None else first):.3e})")
    ref_drag = s.forces()[1] if ref_tol is not None else None
    if at_tol is not None:
        fields, lift, drag = at_tol
        wall = wall_at_tol
        run_steps = steps
        steps = min(steps, (conv + chunk - 1) // chunk * chunk)
    else:
        fields = s.point_fields()
        lift, drag = s.forces()
        run_steps = steps
    s.close()
    # a state with vacuum points (f clamped to zero over a region) has not
    # reached the flow's steady state, however still it has become
    if conv is not None and not np.all(fields["nd"] > 0.0):
        conv, bad = None, True
    return dict(steps=conv if conv is not None else steps, converged=conv is not None,
                diverged=bad,
                wall_s=wall * (conv / steps if conv else 1.0), steps_run=run_steps,
                first_change=first, change=np.asarray(hist), fields=fields, lift=lift,
                drag=drag, ref_drag=ref_drag, ref_steps=ref_conv)


def _load(path):
    if path and os.path.exists(path):
        with np.load(path, allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    return None


def _save(path, d):
    tmp = path + ".tmp.npz"
    np.savez(tmp, **d)
    os.replace(tmp, path)


def generate(elems, n, seed, rorder, torder, tol, max_steps, data, time_budget=None,
             log=print, geometry="cylinder", precision=_lib.DGBGK_FP64):
    """Steady states of n drawn cases (FP64, or the FP32-bulk form, whose
    steady fields differ from FP64's by ~1e-6), appended to the npz at
    `data` (resumable: cases already there are skipped)."""
    mesh = mesh_for(elems, geometry)
    wd = meshgen.wall_distance(mesh)
    samples = draw(n, seed, geometry)
    d = _load(data) or dict(mach=np.zeros(0), kn=np.zeros(0), steps=np.zeros(0, np.int64),
                            wall_s=np.zeros(0), drag=np.zeros(0), converged=np.zeros(0, bool),
                            y=np.zeros((0, mesh.npoin, 4), np.float32))
    t0 = time.perf_counter()
    for i, s in enumerate(samples):
        if i < len(d["mach"]):
            continue
        if time_budget and time.perf_counter() - t0 > time_budget:
            log(f"  time budget reached after {i} cases")
            break
        r = run_to_steady(cylinder_case(mesh, s, rorder, torder), tol, max_steps, precision,
                          log=log)
        d["mach"] = np.append(d["mach"], s.mach)
        d["kn"] = np.append(d["kn"], s.kn)
        d["steps"] = np.append(d["steps"], r["steps"])
        d["wall_s"] = np.append(d["wall_s"], r["wall_s"])
        d["drag"] = np.append(d["drag"], r["drag"])
        d["converged"] = np.append(d["converged"], r["converged"])
        d["y"] = np.concatenate([d["y"], fields_scaled(r["fields"], s)[None]])
        _save(data, d)
        log(f"  case {i}: M={s.mach:.3f} Kn={s.kn:.4f} {r['steps']} steps "
            f"{'converged' if r['converged'] else 'NOT converged'} {r['wall_s']:.2f}s "
            f"drag={r['drag']:.6f}")
    del wd
    return d


def train(elems, data, params, steps, width=256, hidden=6, batch=65536, lr=1e-3, seed=0,
          log=print, geometry="cylinder", init=None):
    """Adam on the steady states in `data`; from random weights, or from the
    network at `init` (fine-tuning, e.g. on states of another velocity
    grid)."""
    from .mlp import MLP
    mesh = mesh_for(elems, geometry)
    wd = meshgen.wall_distance(mesh)
    d = _load(data)
    keep = d["converged"]
    samples = [Sample(m, k) for m, k in zip(d["mach"][keep], d["kn"][keep])]
    X = np.concatenate([features(mesh, wd, s, geometry) for s in samples])
    Y = np.concatenate([net_targets(y, s) for y, s in zip(d["y"][keep], samples)])
    log(f"  training set: {int(keep.sum())} steady states, {len(X)} points")
    m = MLP(fin=X.shape[1], width=width, hidden=hidden, fout=Y.shape[1], batch=batch)
    if init:
        with np.load(init, allow_pickle=False) as z:
            if (int(z["width"]), int(z["hidden"])) != (width, hidden):
                raise ValueError(f"{init}: network shape differs from width {width}, "
                                 f"hidden {hidden}")
            m.set_params(z["params"])
    else:
        m.init(seed=seed)
    m.set_data(X, Y)
    losses = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        nstep = min(500, steps - done)
        # step decay: full rate, then 0.3x over the last 30%, 0.1x over the last 10%
        f = 1.0 if done < 0.7 * steps else (0.3 if done < 0.9 * steps else 0.1)
        l = m.train(nstep, lr=lr * f, seed=seed + done)
        losses.extend(l.tolist())
        done += nstep
        log(f"  train step {done}: loss {np.mean(l[-50:]):.3e}")
    train_s = time.perf_counter() - t0
    np.savez(params, params=m.get_params(), width=width, hidden=hidden,
             losses=np.asarray(losses, np.float32), train_s=train_s, npoints=len(X),
             ncases=int(keep.sum()), geometry=geometry)
    return m, np.asarray(losses), train_s


def load_model(params, batch: int = 65536):
    """The trained network, for inference in batches of `batch` points (a
    multiple of 64: each batch is one pass of the GEMM kernels)."""
    from .mlp import MLP
    with np.load(params, allow_pickle=False) as z:
        m = MLP(fin=5, width=int(z["width"]), hidden=int(z["hidden"]), fout=4, batch=batch)
        m.set_params(z["params"])
    return m


RUNS = {"warm64": (_lib.DGBGK_FP64, True), "cold32": (_lib.DGBGK_FP32, False),
        "warm32": (_lib.DGBGK_FP32, True)}


def evaluate(elems, params, n, seed, rorder, torder, tol, max_steps, log=print,
             runs=("warm64", "cold32", "warm32"), ref_tol=None, geometry="cylinder",
             samples=None):
    """Per held-out case (n drawn with seed, or the given samples): the FP64
    run from the freestream (cold64) and the requested others (FP64/FP32-
    bulk, from the freestream or from the network's state), all judged by
    cold64's step-1 change.  With ref_tol, cold64 goes on to change <
    ref_tol x its step-1 change, and every run's drag is also compared with
    that converged drag (drag_rel_ref)."""
    m = load_model(params)
    mesh = mesh_for(elems, geometry)
    wd = meshgen.wall_distance(mesh)
    out = []
    for s in (samples if samples else draw(n, seed, geometry)):
        case = cylinder_case(mesh, s, rorder, torder)
        pred = untarget(m.forward(features(mesh, wd, s, geometry)), s)
        cold64 = run_to_steady(case, tol, max_steps, _lib.DGBGK_FP64, log=log, ref_tol=ref_tol)
        norm = cold64["first_change"]
        done = dict(cold64=cold64)
        for k in runs:
            prec, warm = RUNS[k]
            done[k] = run_to_steady(case, tol, max_steps, prec, pred if warm else None, norm,
                                    log=log)
        ref = cold64["fields"]
        rec = dict(mach=s.mach, kn=s.kn, norm=norm)
        dref = cold64["ref_drag"]
        if dref is not None:
            rec["ref_drag"] = float(dref)
            rec["ref_steps"] = cold64["ref_steps"]
        for k, r in done.items():
            rec[k] = dict(steps=int(r["steps"]), converged=bool(r["converged"]),
                          wall_s=float(r["wall_s"]), drag=float(r["drag"]),
                          drag_rel=float(abs(r["drag"] - cold64["drag"]) /
                                         max(abs(cold64["drag"]), 1e-30)),
                          rho_rel_l2=float(np.linalg.norm(r["fields"]["rho"] - ref["rho"]) /
                                           np.linalg.norm(ref["rho"])))
            if dref is not None:
                rec[k]["drag_rel_ref"] = float(abs(r["drag"] - dref) / max(abs(dref), 1e-30))
        # how close the network's state is to the steady one
        rec["pred_rho_rel_l2"] = float(np.linalg.norm(pred[0] - ref["rho"]) /
                                       np.linalg.norm(ref["rho"]))
        out.append(rec)
        log(f"  held-out M={s.mach:.3f} Kn={s.kn:.4f}: " + ", ".join(
            f"{k} {rec[k]['steps']} steps {rec[k]['wall_s']:.2f}s"
            f"{'' if rec[k]['converged'] else ' (not converged)'}"
            for k in done) + "; drag rel " + ", ".join(
            f"{k} {rec[k]['drag_rel']:.1e}" for k in runs) +
            ("" if dref is None else "; vs converged FP64 drag " + ", ".join(
                f"{k} {rec[k]['drag_rel_ref']:.1e}" for k in done)) +
            f"; prediction rho rel L2 {rec['pred_rho_rel_l2']:.1e}")
    return out


def time_to_residual(case, params, sample: Sample, geometry: str = "cylinder",
                     tol: float = 1e-6, max_steps: int = 2_000_000, ref_tol: float | None = None,
                     log=None, chunk: int = 1000):
    """Steps and wall time to a step change of tol times the plain run's
    step-1 change: FP64 from the freestream (plain) against FP32-bulk from
    the Maxwellian of the network's state (learned), both judged by the
    plain run's step-1 change.  The learned run's wall time includes the
    network's inference.  With ref_tol the plain run goes on to ref_tol for
    a converged reference drag."""
    mesh = case.mesh
    plain = run_to_steady(case, tol, max_steps, _lib.DGBGK_FP64, log=log, ref_tol=ref_tol,
                          chunk=chunk)
    if log:
        log(f"#   plain FP64: {plain['steps']} steps, {plain['wall_s']:.1f} s")
    t0 = time.perf_counter()
    m = load_model(params)
    pred = untarget(m.forward(features(mesh, meshgen.wall_distance(mesh), sample, geometry)),
                    sample)
    m.close()
    predict_s = time.perf_counter() - t0
    learned = run_to_steady(case, tol, max_steps, _lib.DGBGK_FP32, pred, plain["first_change"],
                            log=log, chunk=chunk)
    if log:
        log(f"#   learned FP32-bulk: {learned['steps']} steps, {learned['wall_s']:.1f} s "
            f"(+ network {predict_s:.2f} s)")
    wl = learned["wall_s"] + predict_s
    out = {"tol": tol,
           "plain_fp64": {"steps": int(plain["steps"]), "wall_s": plain["wall_s"],
                          "converged": bool(plain["converged"]), "drag": plain["drag"]},
           "learned_fp32": {"steps": int(learned["steps"]), "wall_s": wl, "network_s": predict_s,
                            "converged": bool(learned["converged"]), "drag": learned["drag"]},
           "speedup_wall": plain["wall_s"] / wl,
           "speedup_steps": plain["steps"] / max(learned["steps"], 1),
           "drag_rel": abs(learned["drag"] - plain["drag"]) / abs(plain["drag"]),
           "pred_rho_rel_l2": float(np.linalg.norm(pred[0] - plain["fields"]["rho"]) /
                                    np.linalg.norm(plain["fields"]["rho"]))}
    if ref_tol is not None and plain["ref_drag"] is not None:
        ref = plain["ref_drag"]
        out["ref_tol"] = ref_tol
        out["ref_drag"] = ref
        out["ref_steps"] = plain["ref_steps"]
        out["plain_fp64"]["drag_rel_ref"] = abs(plain["drag"] - ref) / abs(ref)
        out["learned_fp32"]["drag_rel_ref"] = abs(learned["drag"] - ref) / abs(ref)
    return out


def summarise(res):
    kinds = [k for k in ("cold64", "warm64", "cold32", "warm32") if k in res[0]]

    def med(key, a, b):
        return float(np.median([r[a][key] / max(r[b][key], 1e-30) for r in res]))
    ok = [r for r in res if all(r[k]["converged"] for k in kinds)]
    out = dict(n=len(res), all_converged=len(ok))
    for k in kinds[1:]:
        out[f"speedup_steps_{k}_vs_cold64"] = med("steps", "cold64", k)
        out[f"speedup_wall_{k}_vs_cold64"] = med("wall_s", "cold64", k)
        out[f"median_drag_rel_{k}"] = float(np.median([r[k]["drag_rel"] for r in res]))
        out[f"max_drag_rel_{k}"] = float(max(r[k]["drag_rel"] for r in res))
    if "drag_rel_ref" in res[0]["cold64"]:
        for k in kinds:
            out[f"median_drag_rel_ref_{k}"] = float(np.median([r[k]["drag_rel_ref"] for r in res]))
            out[f"max_drag_rel_ref_{k}"] = float(max(r[k]["drag_rel_ref"] for r in res))
    out["median_pred_rho_rel_l2"] = float(np.median([r["pred_rho_rel_l2"] for r in res]))
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="learned warm start: generate, train, evaluate")
    ap.add_argument("cmd", choices=["generate", "train", "evaluate"])
    ap.add_argument("--elems", type=int, default=4000)
    ap.add_argument("--rorder", type=int, default=8)
    ap.add_argument("--torder", type=int, default=16)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--ref-tol", type=float, default=0.0,
                    help="evaluate: also run cold64 on to this tolerance and compare drags with it")
    ap.add_argument("--max-steps", type=int, default=200000)
    ap.add_argument("--time-budget", type=float, default=0.0)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--data", default="warmstart_train.npz")
    ap.add_argument("--params", default="warmstart_params.npz")
    ap.add_argument("--out", default="warmstart_eval.json")
    ap.add_argument("--geometry", default="cylinder", choices=sorted(GEOMETRIES))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="generate: the form the steady states are run in")
    ap.add_argument("--init", default="",
                    help="train: start from this network's weights (fine-tuning)")
    ap.add_argument("--case", action="append", default=[], metavar="MACH,KN",
                    help="evaluate: this case instead of drawn ones (repeatable)")
    ap.add_argument("--runs", default="warm64,cold32,warm32",
                    help="runs besides the FP64 freestream run, of warm64, cold32, warm32")
    a = ap.parse_args(argv)
    log = lambda *x: print(*x, flush=True)  # noqa: E731
    os.makedirs(os.path.dirname(os.path.abspath(a.data)), exist_ok=True)
    if a.cmd == "generate":
        d = generate(a.elems, a.n, a.seed, a.rorder, a.torder, a.tol, a.max_steps, a.data,
                     a.time_budget or None, log, a.geometry,
                     _lib.DGBGK_FP32 if a.precision == "fp32" else _lib.DGBGK_FP64)
        log(json.dumps(dict(cases=len(d["mach"]), converged=int(d["converged"].sum()),
                            median_steps=float(np.median(d["steps"])),
                            total_wall_s=float(d["wall_s"].sum()))))
    elif a.cmd == "train":
        _, losses, train_s = train(a.elems, a.data, a.params, a.train_steps, batch=a.batch,
                                   lr=a.lr, log=log, geometry=a.geometry,
                                   init=a.init or None)
        log(json.dumps(dict(train_s=train_s, loss_first=float(losses[:20].mean()),
                            loss_last=float(losses[-50:].mean()))))
    else:
        res = evaluate(a.elems, a.params, a.n, a.seed, a.rorder, a.torder, a.tol,
                       a.max_steps, log, runs=tuple(x for x in a.runs.split(",") if x),
                       ref_tol=a.ref_tol or None, geometry=a.geometry,
                       samples=[Sample(*map(float, c.split(","))) for c in a.case])
        summ = summarise(res)
        with open(a.out, "w") as fh:
            json.dump(dict(elems=a.elems, ordinates=a.rorder * a.torder, tol=a.tol,
                           summary=summ, cases=res), fh, indent=1)
        log(json.dumps(summ))


if __name__ == "__main__":
    main()
