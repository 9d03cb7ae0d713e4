"""Build libbdda.so (HIP, gfx950) in-tree: block-dda_py_amd/lib/libbdda.so.

Usage:  python block-dda_py_amd/build_lib.py [--jobs N]

Each csrc/*.hip file is compiled with hipcc for gfx950 and each host-only
csrc/*.cpp file with g++, in parallel; the objects are linked into one shared
library exporting the C ABI declared in include/bdda.h.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libbdda.so")
OBJ_DIR = os.path.join(OUT_DIR, "obj")

ARCH = os.environ.get("BDDA_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
FFT_SOURCES = ["kernels_fft.hip", "fft_len_a.hip", "fft_len_b.hip", "fft_len_c.hip", "fft_len_d.hip"]
SOURCES = FFT_SOURCES + ["kernels_block.hip", "kernels_dla.hip", "kernels_fno.hip", "plan.hip"]
# the FFT kernels fold the additions of their known-zero (pruned) inputs,
# which IEEE signed zeros would otherwise keep
FILE_FLAGS = {src: ["-fno-signed-zeros"] for src in FFT_SOURCES}
# host-only C++ (built with the host compiler)
CXX_SOURCES = ["host_la.cpp"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC"]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = ["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = ["hipcc", *FLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed for {src}:\n{r.stderr}")
    return obj


def build(jobs=8, verbose=True):
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES + CXX_SOURCES))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
