"""Build recipe for the native libraries (in-tree, gfx950 only).

    python -m clifft_amd.build   # clifft_amd/libclifft_svm.so + libclifft_compile.so

The library is compiled with hipcc for gfx950 with packed-FP32 instruction
selection disabled for the device code (-Xclang -target-feature -Xclang
-packed-fp32-ops).  hipcc hands the flag to both compilation passes; the
x86 host pass does not know the feature and prints "'-packed-fp32-ops' is
not a recognized feature for this target (ignoring feature)", while the
gfx950 device pass applies it: a device-only build of float2 fma code gives
v_pk_fma_f32 without the flag and two v_fma_f32 with it.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

LIB_SVM = os.path.join(HERE, "libclifft_svm.so")
SVM_SOURCES = ["svm_kernels.hip", "svm_host.cpp"]
LIB_PROBE = os.path.join(HERE, "libclifft_probe.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-ffp-contract=fast",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_svm(verbose: bool = True) -> str:
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in SVM_SOURCES:
        obj = os.path.join(HERE, "_build", src + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        _run([HIPCC, *HIP_FLAGS, *lang, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB_SVM + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_SVM)
    return LIB_SVM


def build_probe() -> str:
    """Bandwidth probes for bench.py (not part of the executor)."""
    tmp = LIB_PROBE + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "probe.hip")])
    os.replace(tmp, LIB_PROBE)
    return LIB_PROBE


def build_all() -> None:
    build_svm()
    build_probe()
    from clifft_amd import compiler as _compile
    _compile.build()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
