"""Builds libgeowells_hip.so for gfx950 in-tree (next to this file).

The set-up / C-ABI host code is compiled without floating-point contraction
(so the material classification and departure branches reproduce the
reference's FP64 decisions); the kernels are split into the pipe / z-step
unit and one unit per line-kernel family and precision, compiled in parallel.
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


def _run_all(cmds):
    """Runs the commands concurrently; raises if any fails."""
    procs = []
    for cmd in cmds:
        print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])


# host units: set-up, C ABI, grid generators (no floating-point contraction)
HOST_UNITS = ["gw_host.cpp", "gw_grids.cpp"]
KERNEL_UNITS = ["gw_kernels.hip", "gw_line_w64_f64.hip", "gw_line_w64_f32.hip", "gw_line_w32_f64.hip",
                "gw_line_w32_f32.hip", "gw_line_w64_f32sr.hip", "gw_line_w32_f32sr.hip", "gw_line_adj_w64.hip",
                "gw_line_adj_w32.hip", "gw_adjoint.hip", "gw_post.hip"]


def build(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
              "-I", CSRC]
    host_objs = [os.path.join(objdir, u.replace(".cpp", ".o")) for u in HOST_UNITS]
    kern_objs = [os.path.join(objdir, u.replace(".hip", ".o")) for u in KERNEL_UNITS]
    cmds = [[hipcc, *common, "-ffp-contract=off", "-c", os.path.join(CSRC, u), "-o", o]
            for u, o in zip(HOST_UNITS, host_objs)]
    cmds += [[hipcc, *common, *DEVICE_FLAGS, "-c", os.path.join(CSRC, u), "-o", o]
             for u, o in zip(KERNEL_UNITS, kern_objs)]
    _run_all(cmds)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *host_objs, *kern_objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
