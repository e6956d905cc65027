"""Builds the native parts in-tree.

- libdgbgk.so: HIP kernels for gfx950 + the C ABI (include/dgbgk.h), with
  the host setup (prep.cpp) compiled by g++ with -ffp-contract=off so the
  reference's preprocessing arithmetic rounds exactly as the Fortran's.
- dgbgk_run: the Fortran 90 host (deck/mesh readers, output files) that
  drives the library through ISO_C_BINDING, built with amdflang.
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)               # dg-bgk_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
FSRC = os.path.join(ROOT, "fortran")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libdgbgk.so")
FHOST = os.path.join(PKG, "dgbgk_run")
ARCH = os.environ.get("DGBGK_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 instructions are disabled for all device code
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Builds libdgbgk.so (or another build at `out` with extra -D defines)."""
    os.makedirs(BUILD, exist_ok=True)
    prep_o = os.path.join(BUILD, "prep.o")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-c", os.path.join(CSRC, "prep.cpp"), "-o", prep_o])
    objs = [prep_o]
    for src in ("dgbgk.hip", "mlp.hip"):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        tag = "" if not defines else "_" + "_".join(d.replace("=", "") for d in defines)
        o = os.path.join(BUILD, src.replace(".hip", tag + ".o"))
        _run([HIPCC] + HIP_FLAGS + [f"-D{d}" for d in defines] + ["-c", path, "-o", o])
        objs.append(o)
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_fortran_host(verbose: bool = False) -> str | None:
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None or not os.path.isdir(FSRC):
        return None
    srcs = [os.path.join(FSRC, f) for f in ("dgbgk_api.f90", "dgbgk_run.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(BUILD, exist_ok=True)
    tmp = FHOST + ".tmp"
    _run([fc, "-O2", "-J", BUILD] + srcs + ["-o", tmp, "-L" + PKG, "-ldgbgk",
                                            "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FHOST)
    if verbose:
        print("built", FHOST)
    return FHOST


STREAM = os.path.join(PKG, "stream_ceiling")


def build_stream_ceiling(verbose: bool = False) -> str:
    """The streaming-bandwidth probe (csrc/stream_ceiling.hip), an executable."""
    tmp = STREAM + ".tmp"
    _run([HIPCC] + [f for f in HIP_FLAGS if f != "-fPIC"] +
         [os.path.join(CSRC, "stream_ceiling.hip"), "-o", tmp])
    os.replace(tmp, STREAM)
    if verbose:
        print("built", STREAM)
    return STREAM


ACCESS = os.path.join(PKG, "access_probe")


def build_access_probe(verbose: bool = False) -> str:
    """The update kernel's access structure without its arithmetic
    (csrc/access_probe.hip), an executable."""
    tmp = ACCESS + ".tmp"
    _run([HIPCC] + [f for f in HIP_FLAGS if f != "-fPIC"] +
         [os.path.join(CSRC, "access_probe.hip"), "-o", tmp])
    os.replace(tmp, ACCESS)
    if verbose:
        print("built", ACCESS)
    return ACCESS


def build_all(verbose: bool = False):
    build_lib(verbose)
    build_fortran_host(verbose)
    build_stream_ceiling(verbose)
    build_access_probe(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
