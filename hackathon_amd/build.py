"""Build the in-tree gfx950 libraries with hipcc.

    python -m hackathon_amd.build          (or hackathon_amd.build.build())

Produces, next to this file:
  libhh.so       the C-ABI core (include/hh_solver.h): HIP kernels + host API
  libhh_ref2d.so the reference-shaped drop-in solver(System) / forwardDST /
                 reverseDST for the 2D program
  libhh_ref3d.so the drop-in solver(System) / DST for the 3D program
Each .hip file is compiled separately (in parallel) and linked once.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HH_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions (v_pk_*_f32) are disabled: some of their forms
# have produced silently wrong results on these machines
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{INCLUDE}"]

CORE_SOURCES = ["hh_dst.hip", "hh_tri.hip", "hh_stencil.hip", "hh_api.hip", "hh_host.hip", "hh_serve.hip", "hh_modes.hip"]
LIB = os.path.join(HERE, "libhh.so")
REF2D = os.path.join(HERE, "libhh_ref2d.so")
REF3D = os.path.join(HERE, "libhh_ref3d.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "target-feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(CORE_SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, CORE_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib",
          "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    # drop-in reference-shaped front ends (plain C, link against libhh.so)
    for src, out in (("hh_ref2d.c", REF2D), ("hh_ref3d.c", REF3D)):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        tmp = out + ".tmp"
        _run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu99", "-Wall", f"-I{INCLUDE}", "-o", tmp, path,
              f"-L{HERE}", "-lhh", "-Wl,-rpath,$ORIGIN", "-lm"])
        os.replace(tmp, out)
    # Direct_Solver drivers (order N [threads]); -rdynamic exports their
    # user-input functions to the drop-in libraries
    for src, lib, out in (("hh_driver2d.c", "hh_ref2d", "Direct_Solver2d"), ("hh_driver3d.c", "hh_ref3d", "Direct_Solver3d")):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        exe = os.path.join(HERE, out)
        _run(["gcc", "-O2", "-std=gnu99", "-Wall", "-rdynamic", f"-I{INCLUDE}", "-o", exe + ".tmp", path,
              f"-L{HERE}", f"-l{lib}", "-lhh", "-Wl,-rpath,$ORIGIN", "-lm"])
        os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", LIB)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
