"""Builds libwiremom.so (HIP for gfx950) and the wire_mom driver in this directory.

    python praca-magisterska_amd/build.py [--jobs N]

Each source compiles to an object under _build/ in parallel, then one link.
Every HIP compile disables packed FP32 instructions
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host pass warns that
it does not know the feature, which is harmless.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libwiremom.so")
DRIVER = os.path.join(HERE, "wire_mom")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["wm_fill.hip", "wm_lu.hip", "wm_solve.hip", "wm_batch.hip", "wm_gmres.hip", "wm_api.hip", "wm_server.hip",
           "wm_transient.hip", "wm_geometry.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
HIPFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-c"]
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + HIPFLAGS
    cmd += [os.path.join(CSRC, src), "-o", obj] + COMMON
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stderr}")
    if err:
        sys.stderr.write(err + "\n")
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, OUT)
    # the reference program's driver: host C++ only, linked against the library beside it
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "wire_mom.cpp"), "-L", HERE, "-lwiremom", "-Wl,-rpath,$ORIGIN",
           "-o", DRIVER + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    print(build(a.jobs))
