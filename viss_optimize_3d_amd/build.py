"""Build libviss_hip.so (HIP, gfx950) and the host-driver executables in-tree.

Usage: python -m viss_optimize_3d_amd.build [--force]
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libviss_hip.so")
BIN = os.path.join(PKG, "bin")

ARCH = os.environ.get("VISS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Extra device flags (e.g. -DVISS_BP_PIX_PER_LANE=4) for launch-shape experiments.
EXTRA = os.environ.get("VISS_EXTRA_FLAGS", "").split()
# Packed FP32 instructions (v_pk_fma_f32 and the like) are disabled for all
# device code; hipcc's host pass warns that it does not know the feature.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", INCLUDE, "-I", CSRC,
          "-Wall", "-Wno-unused-function"] + DEVICE_FLAGS + EXTRA

SOURCES = ["viss_sky.hip", "viss_pix2ang.hip", "viss_nudft.hip", "viss_weights.hip", "viss_batch.hip", "viss_image.hip", "viss_snapshots.hip"]
HOST_PROGRAMS = {"pix2ang": "host/pix2ang_main.cpp", "vissgen3d": "host/vissgen3d_main.cpp",
                 "vissimage_sum": "host/vissimage_sum_main.cpp"}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build(force: bool = True, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    headers = [os.path.join(CSRC, "viss_internal.h"), os.path.join(CSRC, "viss_kernels.h"), os.path.join(INCLUDE, "viss.h")]
    jobs = []
    objs = []
    for src in SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _newer(o, [s] + headers):
            jobs.append([HIPCC] + COMMON + ["-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(4, max(1, len(jobs)))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    if force or _newer(LIB, objs):
        tmp = LIB + ".tmp"
        _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs)
        os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    for name, src in HOST_PROGRAMS.items():
        s = os.path.join(CSRC, src)
        if not os.path.exists(s):
            continue
        exe = os.path.join(BIN, name)
        if force or _newer(exe, [s, LIB] + headers):
            _run([HIPCC, "-O2", "-std=c++17", "-fopenmp", "-I", INCLUDE, s, "-o", exe + ".tmp", "-L", PKG,
                  "-lviss_hip", "-Wl,-rpath,$ORIGIN/.."])
            os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print(LIB)
