"""Build libpointint_hip.so (gfx950) from pointint_amd/csrc with hipcc.

Every HIP translation unit is compiled for gfx950 only, with packed FP32
instructions disabled, and linked into one shared library next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "libpointint_hip.so")
CXX_DIR = os.path.join(HERE, "cxx")
OUT_CXX = os.path.join(HERE, "libpointint_cxx.so")
OUT_EXAMPLE = os.path.join(HERE, "example_cube_pair")
OUT_ALIGN = os.path.join(HERE, "pointint_align")
ROCM_INCLUDE = "/opt/rocm/include"
ROCM_LIB = "/opt/rocm/lib"
BUILD_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", INCLUDE, "-Wall", "-Wno-unused-function",
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


# Per-file options: the per-face kernel keeps its matrix-core products in
# VGPRs (no AGPR copies inside its face loop).
PER_FILE = {"ff_fast.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, *PER_FILE.get(os.path.basename(src), []), "-c", src,
           "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_cxx()
    if verbose:
        print(f"built {OUT}, {OUT_CXX}, {OUT_EXAMPLE}, {OUT_ALIGN}")
    return OUT


def build_cxx() -> None:
    """The reference-named C++ API (host code on the C ABI: matrix driver,
    kernel registry and alignment) and its two programs, compiled with g++ and
    linked against libpointint_hip.so (and the HIP runtime, for the device
    buffers of caller-registered plug-ins)."""
    cxx = os.environ.get("CXX", "g++")
    common = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE,
              "-I", ROCM_INCLUDE, "-Wall"]

    def run(out, cmd, what):
        tmp = out + ".tmp"
        r = subprocess.run([*cmd, "-o", tmp], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{what} build failed: {' '.join(cmd)}\n{r.stderr}")
        os.replace(tmp, out)

    run(OUT_CXX, [cxx, *common, "-shared", os.path.join(CXX_DIR, "pointint_cxx.cpp"),
                  os.path.join(CXX_DIR, "align.cpp"), "-L", HERE, "-lpointint_hip",
                  "-L", ROCM_LIB, "-lamdhip64", "-Wl,-rpath,$ORIGIN"], "C++ API")
    for out, src in ((OUT_EXAMPLE, "example_cube_pair.cpp"), (OUT_ALIGN, "align_main.cpp")):
        run(out, [cxx, *common, os.path.join(CXX_DIR, src), "-L", HERE, "-lpointint_cxx",
                  "-lpointint_hip", "-L", ROCM_LIB, "-lamdhip64", "-Wl,-rpath,$ORIGIN"], src)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
