"""Build libmaxlq.so in-tree with hipcc for gfx950.

    python -m maxlq_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmaxlq.so")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["mlq_engine_fp64.hip", "mlq_engine_fp32.hip", "mlq_noalloc.hip", "mlq_capi.cpp",
           "mlq_io.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MLQ_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions are disabled for every device pass
DEVFLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wno-unused-result", "-Wno-unused-value",
          "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + CFLAGS + DEVFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + CFLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + DEVFLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {src}")
    return obj


def build(jobs: int = 3, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_fortran(verbose)
    return OUT


FDIR = os.path.join(HERE, "fortran")
FSMOKE = os.path.join(FDIR, "mlq_fsmoke")
FC = os.environ.get("MLQ_FC", "/opt/rocm/bin/amdflang")


def _fc(args, cwd):
    r = subprocess.run([FC] + args, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("Fortran build failed: " + " ".join(args[-4:]))


REF_GLOBALS = "/root/reference/global_variables.f90"


def build_fortran(verbose: bool = True) -> str | None:
    """The Fortran interface module and its callers (amdflang): the smoke
    caller, the B-continuation driver, the fftw3-mpi.f03 stand-in check and,
    where the reference's sources are mounted, the check of the drop-in
    modules against the reference's own global_variables module (compiled
    where it lies; objects stay in the build directory)."""
    if not os.path.exists(FC):
        if verbose:
            print("amdflang not found: Fortran callers not built")
        return None
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    link = ["-L", HERE, "-lmaxlq", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."]
    iface = os.path.join(FDIR, "maxlq_c.f90")
    progs = {"mlq_fsmoke": [iface, os.path.join(FDIR, "mlq_fsmoke.f90")],
             "mlq_continuation": [iface, os.path.join(FDIR, "mlq_continuation.f90")],
             "mlq_fftwmpi_check": [os.path.join(FDIR, "mlq_fftwmpi_check.f90")]}
    for name, srcs in progs.items():
        out = os.path.join(FDIR, name)
        _fc(["-O2", "-J", mod, "-I", FDIR] + srcs + ["-o", out + ".tmp"] + link, mod)
        os.replace(out + ".tmp", out)
        if verbose:
            print(f"built {out}")
    if os.path.exists(REF_GLOBALS):
        rmod = os.path.join(OBJ, "fortran_ref")
        os.makedirs(rmod, exist_ok=True)
        srcs = [REF_GLOBALS, iface, os.path.join(FDIR, "fftwfunction_gpu.f90"),
                os.path.join(FDIR, "mlq_field_ops.f90"), os.path.join(FDIR, "mlq_dropin_check.f90")]
        out = os.path.join(FDIR, "mlq_dropin_check")
        _fc(["-O2", "-J", rmod] + srcs + ["-o", out + ".tmp"] + link, rmod)
        os.replace(out + ".tmp", out)
        if verbose:
            print(f"built {out}")
    return FSMOKE


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=3)
    a = ap.parse_args()
    build(a.jobs)
