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
import re
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
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC",
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


RESOURCES = os.path.join(LIB, "kernel_resources.txt")

# Kernels whose LDS arrays are filled by direct global->LDS loads
# (__builtin_amdgcn_global_load_lds), with the bytes their declared __shared__
# arrays take.  If the optimiser ever treats such an array as unwritten it
# drops it (the kernel's LDS size shrinks) and folds its reads to undefined
# values; the build then stops here instead of shipping wrong scores.
DIRECT_LDS_KERNELS = {
    "k_assign": 2 * (32 * 1024 + 1024),             # ztile[2][32 centroid rows + 32 weights]
    "k_assign_wide": 2 * (32 * 1040 + 256),         # ztile[2][32 rows at a 1040 B pitch + weights]
    "k_leaf_scores": 2048 * 4 + 2 * 32 * 1024,      # mid[2048], tile[2][32 member rows]
    "k_leaf_scores_wide": 2048 * 4 + 2 * 32 * 1040,  # mid[2048], tile[2][32 rows at 1040 B]
}

_MANGLED = re.compile(r"_ZN12_GLOBAL__N_1(\d+)")


def kernel_short_name(mangled: str) -> str:
    """`k_assign_wide` from `_ZN12_GLOBAL__N_113k_assign_wideILb0EE...`."""
    m = _MANGLED.search(mangled)
    if not m:
        return mangled
    return mangled[m.end():m.end() + int(m.group(1))]


def parse_resource_remarks(text: str) -> list[dict]:
    """Per-kernel records from clang's -Rpass-analysis=kernel-resource-usage remarks."""
    out, cur = [], None
    for line in text.splitlines():
        if "remark:" not in line or "kernel-resource-usage" not in line:
            continue
        body = line.split("remark:", 1)[1].split("[-Rpass-analysis", 1)[0].strip()
        if body.startswith("Function Name:"):
            cur = {"mangled": body.split(":", 1)[1].strip()}
            cur["kernel"] = kernel_short_name(cur["mangled"])
            out.append(cur)
        elif cur is not None and ":" in body:
            key, val = body.rsplit(":", 1)
            cur[key.strip()] = val.strip()
    return out


def check_direct_lds(records: list[dict]) -> list[str]:
    """Kernels of DIRECT_LDS_KERNELS whose LDS size differs from their declared arrays."""
    bad, seen = [], set()
    for r in records:
        want = DIRECT_LDS_KERNELS.get(r["kernel"])
        if want is None:
            continue
        seen.add(r["kernel"])
        got = int(r.get("LDS Size [bytes/block]", "-1"))
        if got != want:
            bad.append(f"{r['mangled']}: LDS {got} B, declared arrays {want} B")
    for k in DIRECT_LDS_KERNELS:
        if k not in seen:
            bad.append(f"{k}: no resource record")
    return bad


def build_knng() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libknng.so")
    cmd = [hipcc(), *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-shared", f"-I{INC}",
           os.path.join(CSRC, "knng.hip"), os.path.join(CSRC, "knng_serve.hip"),
           os.path.join(CSRC, "knng_batch.cpp"), "-o", out + ".tmp"]
    print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
    # the remarks (and their source context) are parsed below; warnings and errors are shown
    rest = [l for l in lines if ("warning:" in l or "error:" in l) and "kernel-resource-usage" not in l]
    if r.returncode != 0:
        print("\n".join(l for l in lines if "kernel-resource-usage" not in l))
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]}")
    if any(l.strip() for l in rest):
        print("\n".join(rest))
    records = parse_resource_remarks(r.stdout)
    with open(RESOURCES, "w") as f:
        f.write("# per-kernel resources of libknng.so (clang kernel-resource-usage remarks)\n")
        f.write("# kernel\tVGPRs\tAGPRs\tscratch_B_per_lane\toccupancy_waves_per_SIMD\tLDS_B\tmangled\n")
        for rec in records:
            f.write("\t".join([rec["kernel"], rec.get("VGPRs", "?"), rec.get("AGPRs", "?"),
                               rec.get("ScratchSize [bytes/lane]", "?"),
                               rec.get("Occupancy [waves/SIMD]", "?"),
                               rec.get("LDS Size [bytes/block]", "?"), rec["mangled"]]) + "\n")
    bad = check_direct_lds(records)
    if bad:
        raise RuntimeError("direct-load LDS arrays were dropped by the compiler:\n  " + "\n  ".join(bad))
    os.replace(out + ".tmp", out)
    return out


def build_h5() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libknng_h5.so")
    _run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", f"-I{INC}", f"-I{HDF5_ROOT}/include",
          os.path.join(CSRC, "h5io.cpp"), f"-L{HDF5_ROOT}/lib", "-lhdf5",
          # system libstdc++ must win over the older copy next to libhdf5
          f"-Wl,-rpath,{SYSLIB}", f"-Wl,-rpath,{HDF5_ROOT}/lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cli() -> str:
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "knng_main")
    _run(["g++", "-O2", "-std=c++20", f"-I{INC}", os.path.join(CSRC, "knng_main.cpp"),
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
