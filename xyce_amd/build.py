"""Build libxg_load.so (gfx950 HIP + host C++) in-tree.

    python -m xyce_amd.build            # builds xyce_amd/libxg_load.so

Every source is compiled for gfx950 only, in FP64 with floating-point
contraction off (the load must reproduce the reference's rounding), and with
packed-FP32 instruction selection disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxg_load.so")

SOURCES = [
    "xg_api.hip",
    "xg_mos1.hip",
    "xg_bsim3.hip",
    "xg_reduce.hip",
    "xg_linear.hip",
    "xg_tia.hip",
    "xg_lu.hip",
    "xg_front.hip",
    "xg_tran.hip",
    "xg_lu_setup.cpp",
    "xg_setup.cpp",
    "xg_bsim3_setup.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-I" + os.path.join(ROOT, "include"),
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJDIR, src + ".o")
    cmd = [HIPCC] + COMMON + DEVICE
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
