"""Build libsdg_hip.so in-tree for gfx950 (hipcc) -- no JIT, no site-packages install.

Usage: python -m subrosadg_amd.build
"""
from __future__ import annotations

import os
import shutil
from concurrent.futures import ThreadPoolExecutor
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsdg_hip.so")
BUILD_DIR = os.path.join(HERE, "_build")

HOST_SOURCES = ["kit.cpp", "sdg_mesh_api.cpp", "sdg_raw.cpp", "sdg_vtu.cpp", "sdg_gmsh.cpp"]
# the system zstd runtime (raw state files); only its stable C API is used
LINK_LIBS = ["-l:libzstd.so.1"]
HIP_SOURCES = ["sdg_solver.hip", "sdg_batch.hip"]
EXAMPLES = ["taylor_green_3d_cns", "lid_driven_cavity_2d_cns", "shear_layer_2d_cns_restart", "thermal_cavity_3d_cns",
            "explosion_2d_ceuler", "mesh_file_2d_ceuler", "sod_1d_ceuler",
            "pulsed_inflow_2d_ceuler"]
ARCH = os.environ.get("SDG_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 VALU instructions are disabled for all device code; hipcc's host pass warns that it
# does not know the feature, which is harmless.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    """Compile every source and link libsdg_hip.so; always recompiles."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    hipcc = _hipcc()
    t0 = time.time()
    jobs = [["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD_DIR, src + ".o")]
            for src in HOST_SOURCES]
    jobs += [[hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD_DIR, src + ".o")]
             for src in HIP_SOURCES]
    objs = [cmd[-1] for cmd in jobs]
    # the sources compile independently; the HIP ones dominate, so they start first
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for fut in [pool.submit(_run, cmd, verbose) for cmd in reversed(jobs)]:
            fut.result()
    tmp = OUT + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, *LINK_LIBS], verbose)
    os.replace(tmp, OUT)
    # C++ drop-in driver example (include/subrosadg_amd/SubrosaDG.hpp) linked against the library
    root = os.path.dirname(HERE)
    exe_dir = os.path.join(root, "examples", "bin")
    os.makedirs(exe_dir, exist_ok=True)
    ex_jobs = [["g++", "-std=c++23", "-O2", "-I", os.path.join(root, "include"), os.path.join(root, "examples", ex + ".cpp"),
                "-L", HERE, "-lsdg_hip", "-Wl,-rpath,$ORIGIN/../../subrosadg_amd", "-o", os.path.join(exe_dir, ex)]
               for ex in EXAMPLES]
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for fut in [pool.submit(_run, cmd, verbose) for cmd in ex_jobs]:
            fut.result()
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f}s", flush=True)
    return OUT


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
