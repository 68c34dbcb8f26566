"""Build of the native library ``ablate_amd/libablate_fv.so`` (HIP kernels for gfx950 + host C++).

Host-only translation units are compiled with g++, the HIP unit with hipcc for gfx950, and everything is linked
in-tree so that the built library travels with the source tree.  ``build()`` always recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD_DIR = HERE / "_build"
LIB_PATH = HERE / "libablate_fv.so"
ARCH = "gfx950"

HOST_SOURCES = ["mesh.cpp", "mesh_polytope.cpp", "mesh_gmsh.cpp", "mesh_dmplex.cpp", "fv_setup.cpp"]
HIP_SOURCES = ["fv_solver.hip", "hbm_probe.hip"]

# Packed-FP32 instructions are disabled for all device code (the compiler would otherwise form them from
# ordinary float arithmetic).
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the ablate_amd native library cannot be built")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile the native library for gfx950 (translation units in parallel); returns its path."""
    from concurrent.futures import ThreadPoolExecutor

    BUILD_DIR.mkdir(exist_ok=True)
    hipcc = _hipcc()
    steps, objs = [], []
    for src in HIP_SOURCES:  # the long HIP compiles first
        obj = BUILD_DIR / (src + ".o")
        steps.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    for src in HOST_SOURCES:
        obj = BUILD_DIR / (src + ".o")
        steps.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    with ThreadPoolExecutor(max_workers=min(8, len(steps))) as pool:
        for fut in [pool.submit(_run, cmd, verbose) for cmd in steps]:
            fut.result()
    tmp = LIB_PATH.with_name(LIB_PATH.name + ".tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
