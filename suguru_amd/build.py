"""Build libibx.so, libmemb.so and libfluid.so (gfx950) in-tree with hipcc.

    python -m suguru_amd.build            # or __graft_entry__.build()

Flags: -ffp-contract=off keeps every FP64 product and sum separately rounded (the library's
results are bit-identical to the reference loop order); packed FP32 instructions are
disabled for the whole object (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libibx.so")
SOURCES = [os.path.join(HERE, "csrc", "ibx.hip")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", "ibx_device.hpp"), os.path.join(ROOT, "include", "ibx.h")]
# the capsule-membrane library: FP64 membrane law and the BF16-MFMA surrogate
MEMB_LIB = os.path.join(HERE, "libmemb.so")
MEMB_SOURCES = [os.path.join(HERE, "csrc", "membrane.hip"), os.path.join(HERE, "csrc", "mlp.hip")]
# the flow update: FFT momentum and pressure solves (no FMA contraction, like the transfer)
FLUID_LIB = os.path.join(HERE, "libfluid.so")
FLUID_SOURCES = [os.path.join(HERE, "csrc", "fluid.hip")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IBX_OFFLOAD_ARCH", "gfx950")


def hipcc_cmd(out: str = LIB, sources=None, contract: bool = False):
    """contract=False: no FMA contraction (the transfer library is bit-exact to the reference
    loop order); the membrane library is checked to a tolerance and may contract."""
    return [
        HIPCC,
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=fast" if contract else "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-variable",
        "-o", out,
        *(sources or SOURCES),
    ]


def _compile(cmd, out, verbose):
    tmp = out + ".tmp"
    cmd = [tmp if c == out else c for c in cmd]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)


def build(verbose: bool = True) -> str:
    _compile(hipcc_cmd(LIB), LIB, verbose)
    _compile(hipcc_cmd(MEMB_LIB, MEMB_SOURCES, contract=True), MEMB_LIB, verbose)
    _compile(hipcc_cmd(FLUID_LIB, FLUID_SOURCES), FLUID_LIB, verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
