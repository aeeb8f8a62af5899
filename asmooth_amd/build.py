"""Build the gfx950 library, the C++ driver and the CPU oracle, in place.

    python -m asmooth_amd.build          # everything
    python -m asmooth_amd.build --lib    # only libasmooth.so

Outputs (git-ignored, shipped with the source tree):
    asmooth_amd/libasmooth.so        the C ABI of include/asmooth.h
    asmooth_amd/clumping_tree_gpu    C++ driver with the reference's CLI
    oracle/liboracle.so, oracle/_ref/libasmref.so  (test infrastructure)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libasmooth.so")
DRIVER = os.path.join(PKG, "clumping_tree_gpu")

HIP_SOURCES = ["sort.hip", "tree.hip", "knn.hip", "deposit.hip", "halo.hip", "mlp.hip", "api.hip", "calib.hip", "noalloc.hip"]
HOST_SOURCES = ["host_io.cpp"]

# -ffp-contract=off: no FMA contraction, so every FP32 operation rounds as in
# the reference.  Correctly rounded FP32 division and sqrt are requested
# explicitly.  Packed FP32 instructions are disabled for the device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def build_lib(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        cmds.append([hipcc, *HIP_FLAGS, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        cmds.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str:
    src = os.path.join(PKG, "driver", "clumping_tree_gpu.cpp")
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", src, "-I", os.path.join(ROOT, "include"),
          "-L", PKG, "-lasmooth", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "all"])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="only the device library")
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args(argv)
    build_lib(args.j)
    if not args.lib:
        build_driver()
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
