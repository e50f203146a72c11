"""Builds liblisa.so (HIP for gfx950 + host C++) and the `lisa` CLI in-tree.

    python -m lisa_amd.build        # or lisa_amd.build.build()

Outputs go to lisa_amd/_build/ (git-ignored, shipped with the tree).
FP flags: no FMA contraction, correctly rounded FP32 division and sqrt,
denormals kept -- the FP32 results of the kernels then follow the operation
order written in the sources.  Packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LISA_OFFLOAD_ARCH", "gfx950")

FP_FLAGS = ["-ffp-contract=off", "-fno-fast-math"]
HIP_FLAGS = ["-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", SRC,
          "-Wall", "-Wno-unused-result"]

LIB = os.path.join(OUT, "liblisa.so")
CLI = os.path.join(OUT, "lisa")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose_resources=False, defines=(), out_dir=OUT, cli=True):
    """Compile liblisa.so (and the CLI) into out_dir; `defines` are extra
    -D tuning macros (LISA_STACK_LDS, LISA_PATH_MIN_BLOCKS, LISA_WIDE_WAVES, LISA_MODE_WAVES,
    LISA_EMIT_WAVES, LISA_TILE_WAVES; LISA_PATH_BLOCK must stay 64: a workgroup is one wave)."""
    os.makedirs(out_dir, exist_ok=True)
    dev_o = os.path.join(out_dir, "lisa_device.o")
    cache_o = os.path.join(out_dir, "lisa_cache.o")
    host_o = os.path.join(out_dir, "lisa_host.o")
    lib = os.path.join(out_dir, "liblisa.so")
    extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources else []
    extra += [f"-D{d}" for d in defines]
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, *FP_FLAGS, *HIP_FLAGS, *extra, "-c",
          os.path.join(SRC, "lisa_device.hip"), "-o", dev_o])
    # the network's MFMA accumulators in VGPRs (no AGPR copies in the forward loop)
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, *FP_FLAGS, *HIP_FLAGS, *extra,
          "-mllvm", "-amdgpu-mfma-vgpr-form", "-c", os.path.join(SRC, "lisa_cache.hip"), "-o", cache_o])
    _run(["g++", *COMMON, *FP_FLAGS, "-c", os.path.join(SRC, "lisa_host.cpp"), "-o", host_o])
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", dev_o, cache_o, host_o, "-o", tmp,
          "-lpthread"])
    os.replace(tmp, lib)
    if cli:
        _run(["g++", *COMMON, os.path.join(SRC, "lisa_cli.cpp"), "-o",
              os.path.join(out_dir, "lisa"), "-L", out_dir, "-llisa", "-Wl,-rpath,$ORIGIN"])
    return lib


# Variant of the library with a 2-entry LDS traversal stack: every deeper
# entry goes to the per-lane private overflow array, so a parity test run
# against it exercises that path (the default 20-entry stack never
# overflows on the benchmark scenes).
SMALL_STACK = os.path.join(OUT, "stack2")


def build_small_stack():
    return build(defines=("LISA_STACK_LDS=2",), out_dir=SMALL_STACK, cli=False)


if __name__ == "__main__":
    build(verbose_resources="-v" in sys.argv)
    build_small_stack()
