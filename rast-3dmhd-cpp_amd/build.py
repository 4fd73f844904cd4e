"""Builds libmhd3d.so (HIP, gfx950) in-tree: rast-3dmhd-cpp_amd/lib/libmhd3d.so.

Every source under csrc/ (.hip device + host code, .cpp host code) is compiled
to an object in parallel, then the objects are linked into one library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = [os.path.join(CSRC, f)
           for f in ("mhd3d.hip", "mixed.hip", "mixed_ensemble.hip", "ext_stream.hip", "ext_mixed.hip", "dumpio.cpp",
                     "static_state.cpp")]
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "lib", "obj")
OUT = os.path.join(OUT_DIR, "libmhd3d.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # parity: no contraction into FMA, same rounding as the reference's x86-64 build
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    # the 2-wave one-barrier variants of the FP64 MHD stage kernels (kept for
    # the autotuner) need more registers than 2 waves allow; the compiler
    # builds them for 1 wave and says so for each
    "-Wno-pass-failed",
]
# per-source additions: the mixed form's single-domain FP32 kernels measured
# fastest with the instruction-level-parallelism scheduler (MI355X,
# 512x512x256 MHD stage kernel 3.28 ms against 3.35 ms with the default one);
# the ensemble's FP32 kernels (mixed_ensemble.hip) keep the default
SOURCE_FLAGS = {
    "mixed.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}


def _compile(src, extra, verbose):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, *SOURCE_FLAGS.get(os.path.basename(src), []), *extra, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wno-unused-result", *extra, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False, extra=()):
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, list(extra), verbose), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
