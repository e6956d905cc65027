"""Benchmark: DOF-velocity updates per second of the DG-BGK time march.

Workload (default): the plain FP64 cylinder flow -- 20D x 10D box with a
cylinder of diameter D, ~100k-triangle graded Delaunay mesh, 16 x 32 = 512
polar ordinates, Mach 0.5, Kn 0.05, diffuse isothermal wall, the
reference's in-place velocity sweep.  A step is one full ADVNCE loop body
(wall eta, update of every (node, ordinate), wall pass, moments and
collision frequency, residual, forces, next time step); f is resident in
HBM before the timed region starts.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]

Prints one JSON line.  value = 3*NELEM*VNPNT*K / wall seconds of the K
timed steps, run as the library runs them (replayed HIP graphs).
roofline: the update kernel (k_update) against HBM, with algorithmic bytes
= one read + one write of f per DOF-velocity update, divided by its average
launch time over K more steps launched with HIP events around every kernel
on the library's stream.
cpu_baseline: the FP64 C oracle (one core) on a bounded sample of the same
workload (2 steps).
time_to_residual (cylinder configuration; --no-ttr skips it): steps and wall
time to a step change of 1e-6 of the plain run's step-1 change on the same
geometry held at the freestream on all box sides, FP64 from the freestream
against FP32-bulk from the committed warm-start network's state (about a
minute).
"""
import argparse
import json
import os
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


def time_to_residual():
    """The metric's second half on the headline geometry: steps and wall time
    to a step change of 1e-6 of the plain run's step-1 change, FP64 from the
    freestream (plain) against FP32-bulk from the committed network's state
    (learned), on the 101,753-triangle cylinder (16 x 32 ordinates, freestream
    held on all box sides, as steady runs need).  One held-out case near the
    median of the eight evaluated: Mach 0.82, Kn 0.17."""
    from dgbgk import _lib, meshgen, warmstart

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    params = os.path.join(ROOT, "dg-bgk_amd", "dgbgk", "data", "warmstart_cylinder.npz")
    mesh = warmstart.mesh_for(100_000)
    smp = warmstart.Sample(0.82, 0.17)
    case = warmstart.cylinder_case(mesh, smp, 16, 32)
    log(f"# time to residual: cylinder E={mesh.nelem}, Mach {smp.mach}, Kn {smp.kn}")
    plain = warmstart.run_to_steady(case, 1e-6, 400_000, _lib.DGBGK_FP64, log=log)
    log(f"#   plain FP64: {plain['steps']} steps, {plain['wall_s']:.1f} s")
    t0 = time.perf_counter()
    m = warmstart.load_model(params)
    pred = warmstart.untarget(m.forward(warmstart.features(mesh, meshgen.wall_distance(mesh), smp)),
                              smp)
    predict_s = time.perf_counter() - t0
    learned = warmstart.run_to_steady(case, 1e-6, 400_000, _lib.DGBGK_FP32, pred,
                                      plain["first_change"], log=log)
    log(f"#   learned FP32-bulk: {learned['steps']} steps, {learned['wall_s']:.1f} s")
    wl = learned["wall_s"] + predict_s
    return {"case": f"cylinder {mesh.nelem} triangles x 16x32 ordinates, Mach 0.82, Kn 0.17, "
                    "freestream on all box sides",
            "tol": 1e-6,
            "plain_fp64": {"steps": int(plain["steps"]), "wall_s": plain["wall_s"],
                           "converged": bool(plain["converged"]), "drag": plain["drag"]},
            "learned_fp32": {"steps": int(learned["steps"]), "wall_s": wl,
                             "network_s": predict_s,
                             "converged": bool(learned["converged"]), "drag": learned["drag"]},
            "speedup_wall": plain["wall_s"] / wl,
            "speedup_steps": plain["steps"] / max(learned["steps"], 1),
            "drag_rel": abs(learned["drag"] - plain["drag"]) / abs(plain["drag"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--config", choices=["cylinder", "particle"], default="cylinder",
                    help="cylinder: 100k triangles x 512 ordinates (headline); "
                         "particle: 1M triangles x 2048 ordinates")
    ap.add_argument("--elems", type=int, default=0, help="override the mesh size")
    ap.add_argument("--mesh", default="",
                    help="a mesh file in the reference's .plt format, run with the "
                         "cylinder configuration's ordinates and flow conditions")
    ap.add_argument("--shuffle", type=int, default=-1, metavar="SEED",
                    help="renumber the mesh's elements at random (an input mesh in "
                         "arbitrary element order)")
    ap.add_argument("--vchunk", type=int, default=0, help="ordinates per update block (0 = auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ttr", action="store_true",
                    help="skip the time-to-residual comparison (learned FP32-bulk vs plain "
                         "FP64, about a minute; run with the cylinder configuration only)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run three device-to-device copies of BYTES bytes (the "
                         "reference kernel that scales rocprofv3's FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the time march has no multi-GPU path "
              "(a split of the speed rings over GPUs is designed, not built)",
              file=sys.stderr)
        sys.exit(2)

    from dgbgk import _lib, cases, io, meshgen
    from dgbgk.solver import Solver

    t0 = time.time()
    if args.mesh:
        case = cases.cylinder_flow(target_elems=2000)
        case.mesh = io.read_plt(args.mesh)
        wl = (f"mesh {os.path.basename(args.mesh)} x 16x32 ordinates, Mach 0.5, Kn 0.05 "
              "(cylinder flow conditions)")
        mach, kn = 0.5, 0.05
    elif args.config == "cylinder":
        case = cases.cylinder_flow(target_elems=args.elems or 100_000)
        wl = "cylinder flow, 20Dx10D box, ~100k triangles x 16x32 ordinates, Mach 0.5, Kn 0.05"
        mach, kn = 0.5, 0.05
    else:
        case = cases.particle_drag(target_elems=args.elems or 1_000_000)
        wl = "particle drag, ~1M triangles x 32x64 ordinates, Mach 0.3, Kn 1"
        mach, kn = 0.3, 1.0
    if args.shuffle >= 0:
        case.mesh, _ = meshgen.shuffle_elements(case.mesh, args.shuffle)
        wl += f", elements in random order (seed {args.shuffle})"
    prec = _lib.DGBGK_FP64 if args.precision == "fp64" else _lib.DGBGK_FP32
    s = Solver(case, precision=prec, vchunk=args.vchunk)
    setup_s = time.time() - t0
    E, V = case.mesh.nelem, case.nvel
    dofv = 3 * E * V
    print(f"# workload {args.config} E={E} V={V} ({case.rorder}x{case.torder}) "
          f"{args.precision} setup {setup_s:.1f}s", file=sys.stderr, flush=True)
    print("# launch " + " ".join(f"{k}={v}" for k, v in s.launch_config().items()),
          file=sys.stderr, flush=True)

    s.step(args.warmup, records=False)
    s.sync()
    # timed region: K steps as the library runs them (captured HIP graphs
    # of 64 steps, no host synchronisation between steps)
    t1 = time.perf_counter()
    recs = s.step(args.steps)
    s.sync()
    t2 = time.perf_counter()
    wall = t2 - t1
    # kernel timing: the same K steps launched one kernel at a time with
    # HIP events around every launch on the library's stream
    s.timing(True)
    t3 = time.perf_counter()
    s.step(args.steps, records=False)
    s.sync()
    t4 = time.perf_counter()
    upd_ms, upd_n = s.kernel_time(0)
    names = {0: "update", 1: "wall", 2: "moments", 3: "eta", 4: "final", 5: "points"}
    kern = {names[k]: s.kernel_time(k) for k in names}
    s.timing(False)
    print("# kernel ms/launch: " + ", ".join(
        f"{n}={ms / max(c, 1):.4f}" for n, (ms, c) in kern.items() if c), file=sys.stderr)

    bpv = s.bytes_per_value()
    value = dofv * args.steps / wall
    avg_upd_s = (upd_ms / max(upd_n, 1)) / 1e3
    alg_bytes = 2.0 * bpv * dofv  # one read + one write of f per update
    achieved = alg_bytes / avg_upd_s / 1e9 if upd_n else None
    traffic = (_traffic_from_profiles(f"{args.config}:{E}x{V}", args.precision)
               if not args.mesh and args.shuffle < 0 else None)

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_py import OracleRun
        o = OracleRun(case)
        tc0 = time.perf_counter()
        o.step(args.cpu_steps)
        tc = time.perf_counter() - tc0
        cpu = {"value": dofv * args.cpu_steps / tc, "unit": "DOF-velocity updates/s",
               "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same {args.config} workload "
                         f"(E={E}, V={V}), FP64 C oracle, reference order, {tc:.1f} s"}
        o.close()

    out = {
        "metric": "DOF-velocity updates/s + time-to-residual 1e-6, learned vs plain, 1 MI355X",
        "value": value,
        "unit": "DOF-velocity updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (moments f64)",
        "data": "synthetic (seeded graded Delaunay mesh, freestream Maxwellian start)",
        "config": {"workload": wl,
                   "nelem": E, "npoin": case.mesh.npoin, "ordinates": V,
                   "rorder": case.rorder, "torder": case.torder, "mach": mach, "kn": kn,
                   "order": "in-place velocity sweep", "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic,
                     "kernel": "k_update", "alg_bytes_per_launch": alg_bytes,
                     "avg_launch_ms": avg_upd_s * 1e3},
        "kernels_ms_per_launch": {n: ms / c for n, (ms, c) in kern.items() if c},
        "ms_per_step_event_timed": 1e3 * (t4 - t3) / args.steps,
        "residual": {"first": recs[0]["sumres"], "last": recs[-1]["sumres"],
                     "logres_last": recs[-1]["logres"]},
        "cpu_baseline": cpu,
    }
    if (args.config == "cylinder" and not args.no_ttr and not args.elems and not args.mesh
            and args.shuffle < 0):
        out["time_to_residual"] = time_to_residual()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
