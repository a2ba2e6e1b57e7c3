"""Build recipe of the MI355X kNN-graph builder (no cmake/ninja needed).

Products, all in-tree (git-ignored, shipped with the tree):
  lib/libknng.so      HIP kernels + C-ABI (include/knng.h), hipcc --offload-arch=gfx950
  lib/libknng_h5.so   HDF5 boundary (include/knng_h5.h), g++ against /opt/conda HDF5
  bin/knng_main       drop-in CLI (src/main.cpp's command line and output file)

``python build.py`` builds everything; ``__graft_entry__.build()`` calls
``build_all()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
INC = os.path.join(REPO, "include")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")

HDF5_ROOT = os.environ.get("KNNG_HDF5_ROOT", "/opt/conda")
SYSLIB = "/usr/lib/x86_64-linux-gnu"
ARCH = os.environ.get("KNNG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # keep the compiler away from packed-FP32 VALU forms on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        print(out)
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]}")
    if out.strip():
        print(out)


def hipcc() -> str:
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def build_knng() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libknng.so")
    _run([hipcc(), *HIP_FLAGS, "-shared", f"-I{INC}", os.path.join(CSRC, "knng.hip"),
          os.path.join(CSRC, "knng_serve.hip"), os.path.join(CSRC, "knng_batch.cpp"),
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_h5() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libknng_h5.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{INC}", f"-I{HDF5_ROOT}/include",
          os.path.join(CSRC, "h5io.cpp"), f"-L{HDF5_ROOT}/lib", "-lhdf5",
          # system libstdc++ must win over the older copy next to libhdf5
          f"-Wl,-rpath,{SYSLIB}", f"-Wl,-rpath,{HDF5_ROOT}/lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cli() -> str:
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "knng_main")
    _run(["g++", "-O2", "-std=c++17", f"-I{INC}", os.path.join(CSRC, "knng_main.cpp"),
          f"-L{LIB}", "-lknng", "-lknng_h5", "-Wl,-rpath,$ORIGIN/../lib",
          f"-Wl,-rpath,{SYSLIB}", f"-Wl,-rpath,{HDF5_ROOT}/lib", "-Wl,-rpath,/opt/rocm/lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> None:
    build_knng()
    build_h5()
    build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
