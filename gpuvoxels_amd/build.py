"""Builds the engine library in-tree: gpuvoxels_amd/libvx3h.so (gfx950).

    python -m gpuvoxels_amd.build

HIP translation units are compiled by hipcc for gfx950 with packed FP32
instructions disabled; the kernel units without machine loop-invariant code motion
(KERNEL_FLAGS) and without floating-point contraction (the FP64 unit so that it
follows the CPU engine's rounding, both so that the per-step and resident kernels
give identical results).  Host-only units are
compiled by g++.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(os.path.dirname(HERE), "_build")
LIB = os.path.join(HERE, "libvx3h.so")
WORKER = os.path.join(HERE, "vx3_node_worker")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
# Kernel units: no machine-level loop-invariant code motion.  The multi-step kernel
# calls double-precision sin/cos/acos inside its step loop; hoisting their 64-bit
# polynomial coefficients out of the loop pins them in registers for the whole
# launch and forces spills to scratch (e.g. FP64 512x2x1: 336 B/lane of scratch
# with, none without).
KERNEL_FLAGS = ["-mllvm", "-disable-machine-licm"]

UNITS = [
    # (source, compiler, extra flags)
    # both precisions without floating-point contraction: the per-step and resident kernels
    # then round every operation alike and give identical results on either path
    ("vx3_kernels_f32.hip", "hip", KERNEL_FLAGS + ["-ffp-contract=off"]),
    ("vx3_kernels_f64.hip", "hip", KERNEL_FLAGS + ["-ffp-contract=off"]),
    ("vx3h_engine.cpp", "hip", ["-x", "hip"]),
    ("vx3h_build.cpp", "cxx", []),
    ("vx3h_vxa.cpp", "cxx", []),
    ("vx3h_server.cpp", "cxx", []),
    ("vx3h_tune.cpp", "cxx", []),
    ("vx3h_tile.cpp", "cxx", []),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile(unit):
    src, kind, extra = unit
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if kind == "hip":
        cmd = [HIPCC] + HIP_FLAGS + extra + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + CXX_FLAGS + extra + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u[0]))]
    with ThreadPoolExecutor(max_workers=min(8, len(units))) as ex:
        objs = list(ex.map(_compile, units))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    worker_src = os.path.join(CSRC, "vx3_node_worker.cpp")
    if os.path.exists(worker_src):
        _run(["g++"] + CXX_FLAGS + ["-o", WORKER, worker_src, os.path.join(CSRC, "vx3_manager.cpp"), "-L" + HERE, "-lvx3h",
                                    "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
