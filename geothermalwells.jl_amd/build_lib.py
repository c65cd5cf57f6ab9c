"""Builds libgeowells_hip.so for gfx950 in-tree (next to this file).

Two translation units: the set-up / C-ABI host code (compiled without
floating-point contraction so the material classification and departure
branches reproduce the reference's FP64 decisions) and the kernels.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgeowells_hip.so")
ARCH = "gfx950"

# Packed FP32 VALU ops are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
              "-I", CSRC]
    host_obj = os.path.join(objdir, "gw_host.o")
    kern_obj = os.path.join(objdir, "gw_kernels.o")
    _run([hipcc, *common, "-ffp-contract=off", "-c", os.path.join(CSRC, "gw_host.cpp"), "-o", host_obj])
    _run([hipcc, *common, *DEVICE_FLAGS, "-c", os.path.join(CSRC, "gw_kernels.hip"), "-o", kern_obj])
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, host_obj, kern_obj])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
