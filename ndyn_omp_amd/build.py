"""Build recipe of the native parts: the HIP engine (gfx950), the Fortran driver, and the
test-only checkers under oracle/.  Called by __graft_entry__.build()."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libndyn_gpu.so")
FORTRAN = os.path.join(HERE, "fortran")
BIN = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FP64 expressions round in written order (the reference's); FP32 code uses fmaf
    "-ffp-contract=off",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Werror=unused-value", "-Werror=unused-result",
    # the frame writer's host threads
    "-pthread",
]
SOURCES = ["ndg_kernels.hip", "ndg_engine.hip", "ndg_frames.hip"]


def _run(cmd, cwd=None):
    p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return p


def build_hip():
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *[os.path.join(CSRC, s) for s in SOURCES]])
    os.replace(tmp, LIB)


def build_fortran():
    """Fortran driver (deck reader + output writer) over the C ABI, if amdflang exists."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    src_iface = os.path.join(FORTRAN, "ndg_iface.f90")
    src_drv = os.path.join(FORTRAN, "ndyn_gpu.f90")
    if not fc or not os.path.exists(src_drv):
        return False
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "ndyn_gpu")
    _run([fc, "-O2", "-J", BIN, "-o", out, src_iface, src_drv, f"-L{HERE}", "-lndyn_gpu",
          f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."])
    return True


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref-fast"])


def build_all():
    build_hip()
    build_fortran()
    build_oracle()
