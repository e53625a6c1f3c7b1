"""Build the HIP engine library libmaxwellfdtd_amd.so in-tree for gfx950, and beside it
the HBM probe library libmfd_hbm_probe.so (bandwidth measurements, not the solver).

Usage: python -m maxwellfdtd_amd.build   (or maxwellfdtd_amd.build.build()).
Every source is compiled by hipcc for gfx950 only, with packed-FP32 instruction
selection turned off, and linked into a shared library with a C ABI.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmaxwellfdtd_amd.so")
PROBE_LIB = os.path.join(HERE, "libmfd_hbm_probe.so")
SOURCES = ["kernels3d.hip", "kernels3d_p.hip", "kernels3d_f.hip", "kernels2d.hip", "kernels2d_f.hip", "kernels1d.hip",
           "step_kernels.hip",
           "kernels_state.hip", "energy.hip", "engine.cpp", "batch.cpp"]
PROBE_SOURCES = ["hbm_probe.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=fast",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable"]


# sources that must not have multiplies and adds fused behind their back: the fused step
# writes its fused multiply-adds out, so that a cell's H recomputed by a neighbouring
# tile rounds exactly as the owner's (-ffp-contract=fast would let the backend fuse other
# pairs, differently in each kernel instantiation)
NO_CONTRACT = {"kernels3d_f.hip", "kernels2d_f.hip"}


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    flags = [("-ffp-contract=off" if f == "-ffp-contract=fast" and src in NO_CONTRACT else f) for f in FLAGS]
    cmd = [HIPCC, *flags, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def _link(objs, lib):
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of {os.path.basename(lib)} failed:\n{r.stderr}")
    os.replace(tmp, lib)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES + PROBE_SOURCES
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    _link([objs[s] for s in SOURCES], LIB)
    _link([objs[s] for s in PROBE_SOURCES], PROBE_LIB)
    if verbose:
        print(f"built {LIB} and {os.path.basename(PROBE_LIB)}")
    return LIB


if __name__ == "__main__":
    build()
