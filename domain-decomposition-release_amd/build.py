"""Build libddx.so (HIP engine + C ABI) for gfx950, in-tree, and the host
programs over it: lib/ddx_rtm3d (argfile-driven modelling + RTM driver) and
lib/ddx_hosttool (format readers/writers, no GPU).

Usage: python domain-decomposition-release_amd/build.py [--jobs N]

Flags that matter for results:
  -ffp-contract=off                     one rounding per source operation
  -fno-gpu-flush-denormals-to-zero      keep FP32 subnormals (early seismogram
                                        samples are subnormal)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                        no packed FP32 instructions
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libddx.so")
OBJ_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]
SOURCES = ["kernels.hip", "engine.cpp"]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=2, verbose=True):
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


CXX = os.environ.get("CXX", "g++")
HOST = os.path.join(HERE, "host")
HOST_COMMON = [CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), "-I", HOST,
               os.path.join(HOST, "json_lite.cpp"), os.path.join(HOST, "jarvis_host.cpp")]
RTM3D = os.path.join(OUT_DIR, "ddx_rtm3d")
HOSTTOOL = os.path.join(OUT_DIR, "ddx_hosttool")


def _host_jobs(tool_only):
    jobs = [(HOSTTOOL, HOST_COMMON + [os.path.join(HOST, "hosttool.cpp"), "-o", HOSTTOOL + ".tmp"])]
    if not tool_only:
        jobs.append((RTM3D, HOST_COMMON + [os.path.join(HOST, "rtm3d_main.cpp"), "-L", OUT_DIR, "-lddx",
                                           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", RTM3D + ".tmp"]))
    return jobs


def build_host_tool(verbose=False):
    """Only lib/ddx_hosttool (needs no libddx.so)."""
    build_host(verbose, tool_only=True)


def build_host(verbose=True, tool_only=False):
    """Host C++ (g++): the driver links libddx.so (rpath $ORIGIN), the tool
    links nothing from HIP."""
    os.makedirs(OUT_DIR, exist_ok=True)
    for out, cmd in _host_jobs(tool_only):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)
        if verbose:
            print(f"built {out}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=2)
    a = ap.parse_args()
    try:
        build(a.jobs)
        build_host()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
