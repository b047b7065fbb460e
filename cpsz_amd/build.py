"""Builds libcpsz_hip.so in-tree with hipcc for gfx950.

Every translation unit is compiled with -ffp-contract=off (bit-exact float and
double operation order against the reference), without fast-math, and with
packed FP32 instructions disabled for device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcpsz_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ZSTD_INC = os.environ.get("CPSZ_ZSTD_INC", "/opt/conda/include")
ZSTD_LIB = os.environ.get("CPSZ_ZSTD_LIB", "/usr/lib/x86_64-linux-gnu/libzstd.so.1")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + ZSTD_INC, "-Wno-unused-result"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", LIB + ".tmp"] + objs + [ZSTD_LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    build_dropin_demo()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


ROOT = os.path.dirname(HERE)
DEMO_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_roundtrip.cpp")
DEMO_BIN = os.path.join(ROOT, "tests", "cpp", "dropin_roundtrip")


def build_dropin_demo() -> str:
    """A caller of the reference's C++ API compiled against the drop-in header
    (include/sz_cp_preserve_dropin.hpp) and linked with libcpsz_hip.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), DEMO_SRC, "-o", DEMO_BIN + ".tmp",
           "-L" + HERE, "-lcpsz_hip", "-Wl,-rpath,$ORIGIN/../../cpsz_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in demo build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(DEMO_BIN + ".tmp", DEMO_BIN)
    return DEMO_BIN


if __name__ == "__main__":
    build()
    sys.exit(0)
