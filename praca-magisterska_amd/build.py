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


# Sources whose kernels must live in VGPRs alone: the fill kernels hold no MFMA
# accumulators, so AGPRs in use there are spill space, and neither AGPR spill
# space, VGPR spills nor scratch is accepted for them.  (SGPR spills go to VGPR
# lanes, not memory, and are allowed.)
SPILL_FREE = {"wm_fill.hip": ("k_fill", "k_self_terms")}
_REMARK = "kernel-resource-usage"


def resource_usage(stderr: str) -> dict:
    """{mangled kernel name: {field: int}} from -Rpass-analysis=kernel-resource-usage remarks."""
    out, cur = {}, None
    for line in stderr.splitlines():
        if _REMARK not in line or "remark:" not in line:
            continue
        body = line.split("remark:", 1)[1].split("[-Rpass", 1)[0].strip()
        key, _, val = body.partition(":")
        key, val = key.strip(), val.strip()
        if key == "Function Name":
            cur = out.setdefault(val, {})
        elif cur is not None:
            try:
                cur[key] = int(val.split()[0])
            except (ValueError, IndexError):
                pass
    return out


def check_spill_free(src: str, stderr: str) -> None:
    names = SPILL_FREE.get(src)
    if not names:
        return
    usage = {k: v for k, v in resource_usage(stderr).items() if any(n in k for n in names)}
    if not usage:
        raise RuntimeError(f"{src}: no resource-usage remarks for {names}")
    bad = [(k, v) for k, v in usage.items()
           if v.get("AGPRs", 0) or v.get("ScratchSize [bytes/lane]", 0) or v.get("VGPRs Spill", 0)]
    if bad:
        raise RuntimeError(f"{src}: kernels spill (AGPRs, scratch and VGPR spills must be 0):\n"
                           + "\n".join(f"  {k}: {v}" for k, v in bad))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-c"]
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + HIPFLAGS
    if src in SPILL_FREE:
        cmd += [f"-Rpass-analysis={_REMARK}", "-fno-caret-diagnostics"]
    cmd += [os.path.join(CSRC, src), "-o", obj] + COMMON
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stderr}")
    check_spill_free(src, r.stderr)
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and _REMARK not in l and l.strip())
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
