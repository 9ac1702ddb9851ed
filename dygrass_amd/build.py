"""Build recipe for libdygrass (gfx950) and the two drop-in CLIs.

Every HIP source is compiled with hipcc for gfx950 only; the library is a
plain C-ABI shared object (no torch types) loaded by ctypes, and the CLIs
link against it with an $ORIGIN-relative rpath.  Outputs land inside the
package so they travel with the source tree:

    dygrass_amd/lib/libdygrass.so
    dygrass_amd/bin/incremental_update
    dygrass_amd/bin/decremental_update
    dygrass_amd/lib/libdygrass_synth.so   (workload generators, host only)
    dygrass_amd/lib/libdygrass_kappa.so   (kappa check kernels, host only)

Run as ``python -m dygrass_amd.build`` (or through __graft_entry__.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

# IEEE float32 semantics the walk decisions depend on: correctly rounded
# division (1/w, 1/R), no contraction of R*w_e, no packed fp32 ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]

LIB_SOURCES = [
    "walk_kernels.hip", "graph_store.hip", "device_edits.hip", "incremental.hip", "api.hip",
    "stream.hip", "stream_small.hip",
    "mtx_io.cpp",
]
CLIS = {"incremental_update": "cli_incremental.cpp", "decremental_update": "cli_decremental.cpp"}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    path = os.path.join(SRC, src)
    if src.endswith(".hip"):
        _run([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
    else:
        _run([HIPCC, "-x", "c++", *HOST_FLAGS, "-c", path, "-o", obj])
    return obj


def build(verbose: bool = True) -> None:
    for d in (OBJ, LIB, BIN):
        os.makedirs(d, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(LIB_SOURCES))) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    so = os.path.join(LIB, "libdygrass.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, so)
    for exe, src in CLIS.items():
        out = os.path.join(BIN, exe)
        _run([HIPCC, "-x", "c++", *HOST_FLAGS, os.path.join(SRC, src), "-L" + LIB, "-ldygrass",
              "-Wl,-rpath,$ORIGIN/../lib", "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)
    # workload generators (host only, OpenMP)
    synth = os.path.join(LIB, "libdygrass_synth.so")
    _run([CXX, *HOST_FLAGS, "-O3", "-fopenmp", "-shared", os.path.join(SRC, "synth.cpp"),
          "-o", synth + ".tmp"])
    os.replace(synth + ".tmp", synth)
    # FP64 sparse kernels of the kappa check (host only, OpenMP)
    kap = os.path.join(LIB, "libdygrass_kappa.so")
    _run([CXX, *HOST_FLAGS, "-O3", "-fopenmp", "-shared", os.path.join(SRC, "kappa.cpp"),
          "-o", kap + ".tmp"])
    os.replace(kap + ".tmp", kap)
    if verbose:
        print(f"built {so}, {synth} and {', '.join(CLIS)}")


if __name__ == "__main__":
    build()
