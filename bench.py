"""Benchmark: DOF-velocity updates per second of the DG-BGK time march.

Workload (default): the particle-drag case -- a circular particle of
diameter D in a 200D x 100D box, ~1M-triangle graded Delaunay mesh
(1,005,669 triangles), 32 x 64 = 2048 polar ordinates, Mach 0.3, Kn 1,
diffuse isothermal wall, the reference's in-place velocity sweep, plain
FP64.  A step is one full ADVNCE loop body (wall eta, update of every
(node, ordinate), wall pass, moments and collision frequency, residual,
forces, next time step); f is resident in HBM before the timed region
starts.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]
                  [--config particle|cylinder]

Prints one JSON line.  value = 3*NELEM*VNPNT*K / wall seconds of the K
timed steps, run as the library runs them (replayed HIP graphs).
roofline: the update kernel (k_update) against HBM, with algorithmic bytes
= one read + one write of f per DOF-velocity update, divided by its average
launch time over K more steps launched with HIP events around every kernel
on the library's stream.
Extra keys (default run): "fp32_bulk", the same workload and measurement
in the FP32-bulk form; "cylinder", the ~100k x 512 cylinder (Mach 0.5,
Kn 0.05) in both forms.
cpu_baseline: the FP64 C oracle, threaded over the run's OMP_NUM_THREADS
(the same values bit for bit as its serial order), five steps of a
bounded sample of the workload: the same mesh and flow with 2 x 64
ordinates (DOF-velocity updates per second are per ordinate and element,
so the rate carries over).
accuracy: the oracle's result on that sample checks the library's FP64 and
FP32-bulk forms run on the same sample (relative L2 of rho, (u, v) and T at
the points, relative pressure drag on the current point pressure, largest
per-step relative mass difference).
time_to_residual (--no-ttr skips it): steps and wall time to a step change
of 1e-6 of the plain run's step-1 change, FP64 from the freestream against
FP32-bulk from the committed warm-start network's state, freestream held
on all box sides: the particle box at the workload's velocity grid and
flow on a 25,781-triangle mesh ("time_to_residual"), and the ~100k-triangle
cylinder ("time_to_residual_cylinder"); about a minute each.
  python bench.py --network STEPS   times the warm start's network alone.
  python bench.py --contexts N      the update on N contexts created one after
                                    another (the spread between contexts).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dg-bgk_amd"))

HBM_PEAK_GBS = 8000.0


def _traffic_from_profiles(workload: str, precision: str):
    """HBM bytes per update launch from the committed rocprofv3 --pmc summary
    (profiles/update_kernel_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", "update_kernel_traffic.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes[{workload},{precision}]"
    for line in open(path):
        if line.strip().startswith(key):
            try:
                return float(line.split("=", 1)[1].strip())
            except ValueError:
                return None
    return None


def copy_calibration(nbytes: int):
    """Three plain device-to-device copies of a known size: under rocprofv3
    --pmc, the counters' value for this copy gives the scale from counter
    units to bytes read and written."""
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {nbytes} bytes x 3", file=sys.stderr)


TTR_CASES = {
    # geometry: (mesh triangles, rings, angles, Mach, Kn, network)
    "particle": (23_000, 32, 64, 0.3, 1.0, "warmstart_particle.npz"),
    "cylinder": (100_000, 16, 32, 0.82, 0.17, "warmstart_cylinder.npz"),
}


def time_to_residual(geometry: str):
    """The metric's second half: steps and wall time to a step change of 1e-6
    of the plain run's step-1 change, FP64 from the freestream (plain)
    against FP32-bulk from the committed network's state (learned; its
    inference counts), freestream held on all box sides as steady runs need.
    particle: the workload's box, velocity grid (32 x 64) and flow (Mach 0.3,
    Kn 1) on a 25,781-triangle mesh; the 1M-triangle mesh needs ~5x smaller
    steps, ~1e5-1e6 of them at 30 ms, hours on one GPU.  The network was
    fine-tuned on steady states of this grid; this case is held out.
    cylinder: the 101,753-triangle cylinder, 16 x 32 ordinates, Mach 0.82,
    Kn 0.17 (a held-out case near the median of eight evaluated)."""
    from dgbgk import warmstart
    elems, R, T, mach, kn, net = TTR_CASES[geometry]
    mesh = warmstart.mesh_for(elems, geometry)
    smp = warmstart.Sample(mach, kn)
    case = warmstart.cylinder_case(mesh, smp, R, T)
    log(f"# time to residual: {geometry} E={mesh.nelem} V={R * T}, Mach {mach}, Kn {kn}")
    r = warmstart.time_to_residual(case, os.path.join(ROOT, "dg-bgk_amd", "dgbgk", "data", net),
                                   smp, geometry, tol=1e-6, max_steps=400_000, log=log)
    r["case"] = (f"{geometry} {mesh.nelem} triangles x {R}x{T} ordinates, Mach {mach}, Kn {kn}, "
                 "freestream on all box sides")
    return r


def network_bench(steps: int = 200):
    """The warm start's network on its own: Adam steps at batch 65,536 and
    inference of 2^20 points (5 -> 6 x 256 SiLU -> 4, BF16 MFMA kernels), on
    seeded synthetic data.  Under rocprofv3 --kernel-trace this gives the
    per-kernel times; one step's matrix products are 6 x 331,008 x 65,536
    flop (forward, data and weight gradients)."""
    import numpy as np
    from dgbgk.mlp import MLP
    rng = np.random.default_rng(0)
    n = 1 << 20
    X = rng.standard_normal((n, 5)).astype(np.float32)
    Y = np.tanh(X[:, :4] * 0.7 + X[:, 1:5] * 0.3).astype(np.float32)
    m = MLP(fin=5, width=256, hidden=6, fout=4, batch=65536)
    m.init(0)
    m.set_data(X, Y)
    m.train(20, lr=1e-3, seed=1)
    t0 = time.perf_counter()
    loss = m.train(steps, lr=1e-3, seed=2)
    dt = time.perf_counter() - t0
    flop = 6.0 * (5 * 256 + 5 * 256 * 256 + 256 * 4) * 65536
    t1 = time.perf_counter()
    m.forward(X)
    df = time.perf_counter() - t1
    m.close()
    return {"train_steps": steps, "batch": 65536, "ms_per_train_step": 1e3 * dt / steps,
            "train_tflops": flop * steps / dt / 1e12, "loss_last": float(loss[-1]),
            "forward_points": n, "forward_ms": 1e3 * df,
            "forward_tflops": flop / 3 * (n / 65536) / df / 1e12}


KERNELS = {0: "update", 1: "wall", 2: "moments", 3: "eta", 4: "final", 5: "points"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_case(config: str, elems: int = 0, mesh: str = "", shuffle: int = -1):
    """(case, workload text, mach, kn) of a bench workload."""
    from dgbgk import cases, io, meshgen
    if mesh:
        case = cases.cylinder_flow(target_elems=2000)
        case.mesh = io.read_plt(mesh)
        wl = (f"mesh {os.path.basename(mesh)} x 16x32 ordinates, Mach 0.5, Kn 0.05 "
              "(cylinder flow conditions)")
        mach, kn = 0.5, 0.05
    elif config == "cylinder":
        case = cases.cylinder_flow(target_elems=elems or 100_000)
        wl = "cylinder flow, 20Dx10D box, ~100k triangles x 16x32 ordinates, Mach 0.5, Kn 0.05"
        mach, kn = 0.5, 0.05
    else:
        case = cases.particle_drag(target_elems=elems or 1_000_000)
        wl = ("particle drag, 200Dx100D box, ~1M triangles x 32x64 ordinates, "
              "Mach 0.3, Kn 1")
        mach, kn = 0.3, 1.0
    if shuffle >= 0:
        case.mesh, _ = meshgen.shuffle_elements(case.mesh, shuffle)
        wl += f", elements in random order (seed {shuffle})"
    return case, wl, mach, kn


def measure(case, precision: str, steps: int, warmup: int, vchunk: int = 0,
            traffic_key: str | None = None) -> dict:
    """Times `steps` steps as the library runs them (replayed HIP graphs),
    then the same number launched one kernel at a time with HIP events
    around every launch; frees the device context before returning."""
    from dgbgk import _lib
    from dgbgk.solver import Solver
    prec = _lib.DGBGK_FP64 if precision == "fp64" else _lib.DGBGK_FP32
    t0 = time.time()
    s = Solver(case, precision=prec, vchunk=vchunk)
    E, V = case.mesh.nelem, case.nvel
    dofv = 3 * E * V
    log(f"# {case.name} E={E} V={V} ({case.rorder}x{case.torder}) {precision} "
        f"setup {time.time() - t0:.1f}s")
    log("# launch " + " ".join(f"{k}={v}" for k, v in s.launch_config().items()))
    s.step(warmup, records=False)
    s.sync()
    # timed region: K steps with no host synchronisation between them
    t1 = time.perf_counter()
    recs = s.step(steps)
    s.sync()
    wall = time.perf_counter() - t1
    # kernel timing: the same K steps, HIP events around every launch
    s.timing(True)
    t3 = time.perf_counter()
    s.step(steps, records=False)
    s.sync()
    ev_wall = time.perf_counter() - t3
    kern = {KERNELS[k]: s.kernel_time(k) for k in KERNELS}
    s.timing(False)
    log("# kernel ms/launch: " + ", ".join(
        f"{n}={ms / max(c, 1):.4f}" for n, (ms, c) in kern.items() if c))
    upd_ms, upd_n = kern["update"]
    bpv = s.bytes_per_value()
    lc = s.launch_config()
    s.close()
    avg_upd_s = (upd_ms / max(upd_n, 1)) / 1e3
    alg_bytes = 2.0 * bpv * dofv  # one read + one write of f per update
    achieved = alg_bytes / avg_upd_s / 1e9 if upd_n else None
    traffic = _traffic_from_profiles(traffic_key, precision) if traffic_key else None
    return {
        "value": dofv * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic,
                     "kernel": "k_update", "alg_bytes_per_launch": alg_bytes,
                     "avg_launch_ms": avg_upd_s * 1e3},
        "kernels_ms_per_launch": {n: ms / c for n, (ms, c) in kern.items() if c},
        "ms_per_step_event_timed": 1e3 * ev_wall / steps,
        "step_bytes_alg_per_s": 3.0 * bpv * dofv * steps / wall,
        "launch": lc,
        "residual": {"first": recs[0]["sumres"], "last": recs[-1]["sumres"],
                     "logres_last": recs[-1]["logres"]},
        "mass": {"first": recs[0]["mass"], "last": recs[-1]["mass"]},
    }


def host_threads() -> int:
    """Threads for the CPU baseline: the run's OMP_NUM_THREADS (a host that
    grants this process a share of its cores sets it), else the visible CPUs,
    at most 16."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return max(1, min(16, os.cpu_count() or 1))


def _rel_l2(a, b) -> float:
    import numpy as np
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def sample_accuracy(case, nsteps: int, ref: dict, ref_recs: list) -> dict:
    """Relative L2 errors of the point fields rho, (u, v) and T, the relative
    error of the pressure drag (GETFOR on the current point pressure) and
    the largest per-step relative mass difference of the library's FP64 and
    FP32-bulk forms over nsteps steps of the CPU baseline's sample, against
    the FP64 C oracle's values for the same steps (ref: its point fields
    after the last step; ref_recs: its step records)."""
    import ctypes as C
    import numpy as np
    from dgbgk import _lib
    from dgbgk.solver import Solver
    from oracle_py import lib as oracle_lib
    m = case.mesh
    lo, do = C.c_double(), C.c_double()
    oracle_lib().ora_getfor(m.nboun, np.ascontiguousarray(m.bsido, dtype=np.int32),
                            np.ascontiguousarray(ref["ps"], dtype=np.float64),
                            np.ascontiguousarray(m.coord, dtype=np.float64),
                            C.byref(lo), C.byref(do))
    ref_drag = do.value
    out = {"vs": f"FP64 C oracle, {nsteps} steps of the cpu_baseline sample",
           "drag_oracle": ref_drag}
    du = np.hypot(ref["uvel"], ref["vvel"])
    for prec, name in ((_lib.DGBGK_FP64, "fp64"), (_lib.DGBGK_FP32, "fp32_bulk")):
        with Solver(case, precision=prec) as s:
            recs = s.step(nsteps)
            pf = s.point_fields()
            drag = s.forces()[1]
        duv = np.hypot(pf["uvel"] - ref["uvel"], pf["vvel"] - ref["vvel"])
        out[name] = {"rho": _rel_l2(pf["rho"], ref["rho"]),
                     "uv": float(np.linalg.norm(duv) / max(np.linalg.norm(du), 1e-300)),
                     "temp": _rel_l2(pf["temp"], ref["temp"]),
                     "drag": abs(drag - ref_drag) / max(abs(ref_drag), 1e-300),
                     "mass_per_step": max(abs(a["mass"] - b["mass"]) / abs(b["mass"])
                                          for a, b in zip(recs, ref_recs))}
    return out


def cpu_baseline(config: str, elems: int, serial: bool) -> dict:
    """The FP64 C oracle on a bounded sample of the workload: five steps of
    the same mesh and flow with 2 speed rings x 64 angles (Gauss-Legendre
    radial nodes; about 10 s on 16 threads at 1M triangles), threaded over
    the host threads (bit-identical to one thread); optionally also one step
    on one thread.  The oracle's result then checks the library's two forms
    on the same sample ("accuracy")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from dgbgk import cases
    from oracle_py import OracleRun
    if config == "cylinder":
        case = cases.cylinder_flow(target_elems=elems or 100_000, rorder=2, torder=64)
    else:
        case = cases.particle_drag(target_elems=elems or 1_000_000, rorder=2, torder=64)
    o = OracleRun(case)
    dofv = 3 * case.mesh.nelem * case.nvel
    nt = host_threads()
    o.set_threads(nt)
    nsteps = 5
    t0 = time.perf_counter()
    recs = o.step(nsteps)
    dt = (time.perf_counter() - t0) / nsteps
    accuracy = sample_accuracy(case, nsteps, o.point_fields(), recs)
    cpu = None
    for line in open("/proc/cpuinfo"):
        if line.startswith("model name"):
            cpu = line.split(":", 1)[1].strip()
            break
    out = {"value": dofv / dt, "unit": "DOF-velocity updates/s", "cores": nt, "kind": "port",
           "sample": f"{nsteps} steps of the {config} mesh (E={case.mesh.nelem}) and flow with "
                     f"2 x 64 ordinates ({dofv:.3e} DOF-velocity updates per step), FP64 C "
                     f"oracle in the reference's order, {nt} threads, {dt * nsteps:.1f} s",
           "host": {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
                    "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
                    "cpu": cpu or platform.processor()},
           "accuracy": accuracy}
    if serial:
        o.set_threads(1)
        t0 = time.perf_counter()
        o.step(1)
        ds = time.perf_counter() - t0
        out["serial"] = {"value": dofv / ds, "cores": 1, "seconds": ds}
    o.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--config", choices=["particle", "cylinder"], default="particle",
                    help="particle: 1M triangles x 2048 ordinates (headline); "
                         "cylinder: 100k triangles x 512 ordinates")
    ap.add_argument("--elems", type=int, default=0, help="override the mesh size")
    ap.add_argument("--mesh", default="",
                    help="a mesh file in the reference's .plt format, run with the "
                         "cylinder configuration's ordinates and flow conditions")
    ap.add_argument("--shuffle", type=int, default=-1, metavar="SEED",
                    help="renumber the mesh's elements at random (an input mesh in "
                         "arbitrary element order)")
    ap.add_argument("--vchunk", type=int, default=0, help="ordinates per update block (0 = auto)")
    ap.add_argument("--only", action="store_true",
                    help="measure the named configuration and precision only (no extra keys)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-serial", action="store_true",
                    help="also time the CPU baseline sample on one thread")
    ap.add_argument("--no-ttr", action="store_true",
                    help="skip the time-to-residual comparison (learned FP32-bulk vs plain "
                         "FP64 on the cylinder geometry, about a minute)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run three device-to-device copies of BYTES bytes (the "
                         "reference kernel that scales rocprofv3's FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--contexts", type=int, default=0, metavar="N",
                    help="only measure the named configuration and precision on N device "
                         "contexts created one after another (the spread between contexts)")
    ap.add_argument("--network", type=int, default=0, metavar="STEPS",
                    help="only time the warm start's network: STEPS training steps and one "
                         "inference pass (for rocprofv3 kernel traces)")
    args = ap.parse_args()
    if args.network:
        print(json.dumps(network_bench(args.network)))
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the time march has no multi-GPU path "
              "(a split of the speed rings over GPUs is designed, not built)",
              file=sys.stderr)
        sys.exit(2)

    case, wl, mach, kn = make_case(args.config, args.elems, args.mesh, args.shuffle)
    E, V = case.mesh.nelem, case.nvel
    if args.contexts:
        runs = [measure(case, args.precision, args.steps, args.warmup, args.vchunk)
                for _ in range(args.contexts)]
        print(json.dumps({"workload": wl, "precision": args.precision, "contexts": [
            {"update_ms": r["roofline"]["avg_launch_ms"], "frac": r["roofline"]["frac"],
             "ms_per_step": r["ms_per_step"]} for r in runs]}))
        return
    plain = not args.mesh and args.shuffle < 0
    key = f"{args.config}:{E}x{V}" if plain else None
    main_m = measure(case, args.precision, args.steps, args.warmup, args.vchunk, key)

    extra = {}
    if not args.only and plain and not args.elems:
        other = "fp32" if args.precision == "fp64" else "fp64"
        extra["fp32_bulk" if other == "fp32" else "fp64_plain"] = measure(
            case, other, args.steps, args.warmup, args.vchunk, key)
        if args.config == "particle":
            cyl, cwl, _, _ = make_case("cylinder")
            ck = f"cylinder:{cyl.mesh.nelem}x{cyl.nvel}"
            extra["cylinder"] = {"workload": cwl, "nelem": cyl.mesh.nelem,
                                 "ordinates": cyl.nvel,
                                 "fp64": measure(cyl, "fp64", 100, 10, 0, ck),
                                 "fp32_bulk": measure(cyl, "fp32", 100, 10, 0, ck)}
            del cyl

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.config, args.elems, args.cpu_serial)

    out = {
        "metric": "DOF-velocity updates/s + time-to-residual 1e-6, learned vs plain, 1 MI355X",
        "value": main_m["value"],
        "unit": "DOF-velocity updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (moments f64)",
        "data": "synthetic (seeded graded Delaunay mesh, freestream Maxwellian start)",
        "config": {"workload": wl,
                   "nelem": E, "npoin": case.mesh.npoin, "ordinates": V,
                   "rorder": case.rorder, "torder": case.torder, "mach": mach, "kn": kn,
                   "order": "in-place velocity sweep", "precision": args.precision},
        "roofline": main_m["roofline"],
        "kernels_ms_per_launch": main_m["kernels_ms_per_launch"],
        "ms_per_step_event_timed": main_m["ms_per_step_event_timed"],
        "launch": main_m["launch"],
        "residual": main_m["residual"],
        "mass": main_m["mass"],
        "cpu_baseline": cpu,
        "accuracy": cpu.pop("accuracy") if cpu else None,
    }
    out.update(extra)
    if not args.no_ttr and not args.only and plain and not args.elems:
        out["time_to_residual"] = time_to_residual("particle")
        out["time_to_residual_cylinder"] = time_to_residual("cylinder")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
