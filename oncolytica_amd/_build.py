"""Builds liboncohip.so (gfx950) in-tree with hipcc.

The library is the product's compiled core: the fixed kernels
(csrc/onc_kernels.hip) and the C ABI (csrc/onc_runtime.hip, include/onco.h).
csrc/onc_device.h is also embedded as a string so that hiprtc can compile rule
modules against the same device definitions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liboncohip.so")

# Flags: IEEE binary32 semantics identical to the oracle (no contraction,
# correctly rounded divide/sqrt, denormals kept) and packed-FP32 instructions
# disabled.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-value",
    "-Wno-unused-result",
]


def _gen_device_src() -> str:
    src = open(os.path.join(CSRC, "onc_device.h")).read()
    out = os.path.join(CSRC, "onc_device_src.h")
    text = (
        "// Generated by oncolytica_amd/_build.py from onc_device.h; do not edit.\n"
        "#pragma once\n"
        'static const char onc_device_h_src[] = R"ONCDEVSRC(' + src + ')ONCDEVSRC";\n'
    )
    if not os.path.exists(out) or open(out).read() != text:
        with open(out, "w") as f:
            f.write(text)
    return out


def build(verbose: bool = False) -> str:
    """Compile liboncohip.so for gfx950; returns its path."""
    _gen_device_src()
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, "onc_kernels.hip"), os.path.join(CSRC, "onc_runtime.hip")]
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs,
           "-o", tmp, "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
