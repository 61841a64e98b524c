"""Build the native libraries of the HaploADMIXTURE MI355X engine in-tree.

* ``lib/libhaplo_hip.so`` -- hipcc, gfx950: the C ABI declared in ``include/haplo_hip.h``.
* ``lib/libhaplo_sim.so`` -- gcc + OpenMP: the seeded synthetic PLINK BED generator.

Usage: ``python haploadmixture.jl_amd/build.py [--jobs N]``.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)

# every HIP translation unit of csrc/ goes into the one library
HIP_SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 VALU forms are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r.stdout


# translation units with flags of their own: the log-likelihood pass writes its matrix-instruction results to
# VGPRs (its per-pair arithmetic reads every value once; accumulation registers would need a move per value)
EXTRA_FLAGS = {"hk_ll.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def build_hip(jobs: int = 4) -> str:
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), "-I", os.path.join(ROOT, "include"), "-c",
                     os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIB, "libhaplo_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_sim() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libhaplo_sim.so")
    tmp = out + ".tmp"
    cc = shutil.which("gcc") or "cc"
    _run([cc, "-O2", "-fopenmp", "-fPIC", "-shared", "-std=gnu11", os.path.join(CSRC, "hk_sim.c"), "-o", tmp, "-lm"])
    os.replace(tmp, out)
    return out


def build_all(jobs: int = 4):
    return build_sim(), build_hip(jobs)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    for p in build_all(a.jobs):
        print(p)
