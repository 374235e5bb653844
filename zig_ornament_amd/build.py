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


# The sanitizer build (build(sanitize=True), `python build.py --sanitize`):
# the same library with AddressSanitizer and UBSan on its host code only
# (every sanitizer flag applies to the host pass; the GPU code is built as
# usual), for the CPU test suite. It links clang's shared ASan runtime, which
# the process must preload; select the build with ORN_LIB (see _native.py).
SANITIZE_LIB = os.path.join(BUILD, "asan", "libornament_hip.so")
HOST_SANITIZERS = [
    "-Xarch_host", "-fsanitize=address",
    "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-sanitize-recover=undefined",
    "-Xarch_host", "-fno-omit-frame-pointer",
]
SANITIZE_LINK = HOST_SANITIZERS[:4] + ["-Xarch_host", "-shared-libasan"]


def build(verbose=True, sanitize=False):
    out_dir = os.path.join(BUILD, "asan") if sanitize else BUILD
    lib = SANITIZE_LIB if sanitize else LIB
    os.makedirs(out_dir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    common = COMMON + (["-g", *HOST_SANITIZERS] if sanitize else [])
    objs = []
    for src in ("orn_runtime.hip", "orn_nrc.hip"):
        o = os.path.join(out_dir, src.replace(".hip", ".o"))
        _run([HIPCC, *common, *DEVICE, *inc, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    host_o = os.path.join(out_dir, "orn_host.o")
    _run([HIPCC, *common, *inc, "-c", os.path.join(CSRC, "orn_host.cpp"), "-o", host_o])
    tmp = lib + ".tmp"
    link = SANITIZE_LINK if sanitize else []
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", *link, *objs, host_o, "-o", tmp])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build(sanitize="--sanitize" in sys.argv[1:])
    sys.exit(0)
