"""Builds the gfx950 shared library ``sphere_amd/libsphere_dem.so`` in-tree.

``python -m sphere_amd.build`` compiles every source under ``sphere_amd/csrc``
with hipcc (device code for gfx950 only) and links one C-ABI library.
FP contraction is off (the reference builds with --fmad=false,
src/CMakeLists.txt:17) so that the FP64 path reproduces the CPU oracle bit
for bit, and packed-FP32 instruction selection is disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsphere_dem.so")
CLI = os.path.join(HERE, "sphere")
ARCH = os.environ.get("SPHERE_AMD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["kernels.hip", "engine.hip", "adjoint.hip", "analysis.hip", "binio.cpp", "generate.cpp"]
# command-line front ends, each linked against the library next to it
TOOLS = {"sphere": "sphere_main.cpp", "porosity": "porosity_main.cpp", "forcechains": "forcechains_main.cpp",
         "sphere_status": "sphere_status_main.cpp"}


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip", *DEVICE]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    for tool, src in TOOLS.items():
        exe = os.path.join(HERE, tool)
        cmd = [HIPCC, *COMMON, "-o", exe + ".tmp", os.path.join(CSRC, src),
               "-L" + HERE, "-lsphere_dem", "-Wl,-rpath,$ORIGIN"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"{tool} build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {LIB} and {', '.join(os.path.join(HERE, t) for t in TOOLS)}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
