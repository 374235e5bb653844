"""Builds libornament_hip.so (gfx950 kernels + host layer) in-tree with hipcc.

The library must be rebuilt after any change under csrc/ or include/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libornament_hip.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every floating-point expression is rounded as written; the
# kernels place their fused multiply-adds explicitly (reproducible results).
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    objs = []
    for src in ("orn_runtime.hip", "orn_nrc.hip"):
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *COMMON, *DEVICE, *inc, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    host_o = os.path.join(BUILD, "orn_host.o")
    _run([HIPCC, *COMMON, *inc, "-c", os.path.join(CSRC, "orn_host.cpp"), "-o", host_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", *objs, host_o, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
