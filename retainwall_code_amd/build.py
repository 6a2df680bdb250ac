"""Build recipe: libretainwall.so (hipcc, gfx950) and the host CLI.

Every HIP translation unit is compiled with
  --offload-arch=gfx950 -O3 -ffp-contract=off
  -fhip-fp32-correctly-rounded-divide-sqrt
  -Xclang -target-feature -Xclang -packed-fp32-ops
so that fused multiply-adds appear only where the source writes fmaf(),
divisions and square roots are correctly rounded, and no packed-FP32
instructions are emitted.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libretainwall.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LQ_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["lq_physics.hip", "lq_grid.hip", "lq_interact.hip", "lq_pairlist.hip", "lq_columns.hip",
               "lq_surrogate.hip", "lq_capi.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", INCLUDE, "-I", CSRC,
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:4]) + " ...")
    return r


def _newer(src_list, target):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in src_list)


def build_library(force: bool = True, jobs: int = 4) -> str:
    """Compile every HIP TU for gfx950 and link libretainwall.so."""
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    headers.append(os.path.join(INCLUDE, "retainwall.h"))
    objs, procs = [], []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _newer([s] + headers, o):
            procs.append((src, subprocess.Popen([HIPCC, *HIP_FLAGS, "-c", s, "-o", o],
                                                stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    failed = []
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(out + err)
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_host(force: bool = True) -> list:
    """Build the C++ host (LoquatSph/LoquatIo/LoquatGrid + CLI) if present."""
    outs = []
    if not os.path.isdir(HOST):
        return outs
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return outs
    exe = os.path.join(HERE, "loquat")
    if force or _newer(srcs + [LIB], exe):
        _run(["g++", "-O2", "-std=c++17", "-I", INCLUDE, "-I", HOST, *srcs, "-o", exe + ".tmp",
              "-L", HERE, "-lretainwall", "-Wl,-rpath,$ORIGIN"])
        os.replace(exe + ".tmp", exe)
    outs.append(exe)
    return outs


if __name__ == "__main__":
    print(build_library())
    print(build_host())
