"""Build libff_amd.so in-tree with hipcc for gfx950.

Every translation unit is compiled with -ffp-contract=off and IEEE float
division and square root, so the device arithmetic of the intersection test,
the emission frame and the tally normalisation rounds exactly like the
reference's float code. Packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libff_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-ffp-contract=off",
    "-I", INCLUDE, "-I", CSRC,
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    out = []
    for d, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(d, f))
    return sorted(out)


def _compile(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "__")
    obj = os.path.join(BUILD, rel + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip"):
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs, failed = [], []
    with ThreadPoolExecutor(jobs) as ex:
        for src, obj, r in ex.map(_compile, srcs):
            msg = (r.stdout + r.stderr).strip()
            # hipcc's host pass does not know the packed-fp32 feature: harmless
            msg = "\n".join(l for l in msg.splitlines()
                            if "packed-fp32-ops" not in l and "feature flag" not in l)
            if r.returncode != 0:
                failed.append(src)
                print(msg, file=sys.stderr)
            elif verbose and msg:
                print(msg, file=sys.stderr)
            objs.append(obj)
    if failed:
        raise RuntimeError("hipcc failed on: " + ", ".join(os.path.basename(f) for f in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("link of libff_amd.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
