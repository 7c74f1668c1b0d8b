"""Build recipe of the HIP library nans_projekat_amd/libnans_gpu.so (gfx950).

Flags that the arithmetic depends on:
  -ffp-contract=off                      no FMA contraction (the reference's x86-64
                                         SSE build has none, code/build.sh:5-14)
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE-correct binary32 '/' and sqrt
  -fno-gpu-flush-denormals-to-zero       keep binary32 subnormals
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                         no packed-FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnans_gpu.so")
DRIVER_SRC = os.path.join(HERE, "driver", "nans_headless.cpp")
DRIVER = os.path.join(HERE, "nans_headless")
SIM_SRC = os.path.join(HERE, "driver", "nans_sim.cpp")
SIM = os.path.join(HERE, "libnans_sim.so")
OBJ_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), res.stderr))
    return obj


def build(verbose: bool = False) -> str:
    """Compile every csrc/*.hip for gfx950 and link libnans_gpu.so in-tree."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print("built", LIB, DRIVER, "and", SIM)
    return LIB


def build_driver() -> str:
    """The headless driver and the SimUpdateAndRender adapter libnans_sim.so (host
    C++ only; both load libnans_gpu.so with dlopen).  The adapter's camera and
    input arithmetic is binary32 without contraction, like the reference's."""
    cmds = [["g++", "-O2", "-std=c++17", "-Wall", DRIVER_SRC, "-ldl", "-o", DRIVER],
            ["g++", "-O2", "-std=c++17", "-Wall", "-ffp-contract=off", "-shared", "-fPIC", SIM_SRC, "-ldl", "-lm",
             "-o", SIM + ".tmp"]]
    for cmd in cmds:
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("driver build failed:\n%s\n%s" % (" ".join(cmd), res.stderr))
    os.replace(SIM + ".tmp", SIM)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
