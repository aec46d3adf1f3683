"""Build libcrava_hip.so in-tree with hipcc for gfx950.

    python -m crava_amd.build            # or crava_amd.build.build()
    python -m crava_amd.build --incremental

Objects go to crava_amd/_build/ (git-ignored); the shared library is
crava_amd/libcrava_hip.so.  build() recompiles every source; --incremental
skips objects newer than their source and every header.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcrava_hip.so")
HOST_LIB = os.path.join(HERE, "libcrava_host.so")
HOST_SRC = os.path.join(HERE, "host")
HOST_TEST_SRC = os.path.join(HERE, "..", "tests", "cpp", "host_mirror_main.cpp")
HOST_TEST_BIN = os.path.join(HERE, "..", "tests", "cpp", "host_mirror_main")
ARCH = os.environ.get("CRAVA_HIP_ARCH", "gfx950")

# source -> extra flags.  The facies and grid kernels must reproduce
# float/double roundings bit for bit, so contraction into FMA is off there.
SOURCES = {
    "capi.hip": [],
    "fft_static.hip": [],
    "fft_static_ext0.hip": [],
    "fft_static_ext1.hip": [],
    "fft_static_ext2.hip": [],
    "fft_static_ext3.hip": [],
    "fft_static_ext4.hip": [],
    "fft_static_ext5.hip": [],
    "fft.hip": [],
    "posterior.hip": [],
    "tables.cpp": [],
    "kde.cpp": ["-ffp-contract=off"],
    "simulate.hip": [],
    "facies.hip": ["-ffp-contract=off"],
    "nll.hip": [],
    "batch.hip": [],
    "grid.hip": ["-ffp-contract=off"],
    # contraction off by a pragma in the file's own code, on in the included
    # Stockham passes
    "forward.hip": ["-ffp-contract=fast-honor-pragmas"],
    "swdiv.hip": ["-ffp-contract=fast-honor-pragmas"],
    "localnoise.hip": ["-ffp-contract=off"],
    "eig3.cpp": ["-ffp-contract=off"],
}

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I" + os.path.join(HERE, "..", "include"),
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _newest_dependency(src: str) -> float:
    deps = [os.path.join(CSRC, src), os.path.abspath(__file__),
            os.path.join(HERE, "..", "include", "crava_hip.h")]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    return max(os.path.getmtime(d) for d in deps if os.path.exists(d))


def _compile(src: str, extra: list[str], incremental: bool = False) -> str:
    obj = os.path.join(BUILD, src + ".o")
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) > _newest_dependency(src):
        return obj
    cmd = [_hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, incremental: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(lambda s: _compile(s, SOURCES[s], incremental), srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> None:
    """The C++ host mirror (plain g++, over the C ABI only) and its test driver."""
    gxx = shutil.which("g++") or "g++"
    srcs = sorted(os.path.join(HOST_SRC, f) for f in os.listdir(HOST_SRC) if f.endswith(".cpp"))
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(HERE, "..", "include")]
    cmd = [gxx, *flags, "-shared", "-o", HOST_LIB + ".tmp", *srcs, "-L" + HERE, "-lcrava_hip",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host mirror build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    if os.path.exists(HOST_TEST_SRC):
        cmd = [gxx, *flags, "-o", HOST_TEST_BIN, HOST_TEST_SRC, "-L" + HERE, "-lcrava_host", "-lcrava_hip",
               "-Wl,-rpath,$ORIGIN/../../crava_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host mirror test driver build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", HOST_LIB)


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
