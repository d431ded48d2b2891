"""Builds libkomd_hip.so (gfx950) in-tree with hipcc, and the C++ host driver komd_md
(komd_amd/host, g++) linked against it.

Usage: python -m komd_amd.build   (or komd_amd.build.build() from Python)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libkomd_hip.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["komd_api.hip", "k_container.hip", "k_cell.hip", "k_bonded.hip", "k_surrogate.hip", "k_sensors.hip",
           "k_table.hip"]
HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["komd_host.cpp", "main.cpp"]
HOST_EXE = os.path.join(HERE, "komd_md")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KOMD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, OUT)
    build_host()
    if verbose:
        print(f"built {OUT} and {HOST_EXE}")
    return OUT


def build_host() -> str:
    """g++ build of the host driver; it finds libkomd_hip.so next to itself (rpath $ORIGIN)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", *[os.path.join(HOST, s) for s in HOST_SOURCES],
           "-o", HOST_EXE + ".tmp", "-L" + HERE, "-l:libkomd_hip.so", "-Wl,-rpath,$ORIGIN",
           "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host build failed:\n{res.stderr}")
    os.replace(HOST_EXE + ".tmp", HOST_EXE)
    return HOST_EXE


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
