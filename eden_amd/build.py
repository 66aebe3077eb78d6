"""Build the native parts in-tree.

* ``eden_amd/libedenhip.so`` -- the HIP library (kernels + C-ABI), gfx950 only.
* ``oracle/liboracle.so``    -- the CPU restatement (test infrastructure; built by the
  oracle's own Makefile, invoked from here so one call builds everything).

Run ``python -m eden_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "eden_amd", "csrc")
LIB = os.path.join(ROOT, "eden_amd", "libedenhip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]

SOURCES = ["kernels.hip", "network.hip", "edenhip.hip", "cable_setup.cpp", "diag.hip", "adjoint.hip", "output.cpp"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_hip(verbose_resources=False):
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_" + os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose_resources:
            cmd.insert(1, "-Rpass-analysis=kernel-resource-usage")
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_oracle()
    build_hip(verbose_resources="--resources" in sys.argv)


if __name__ == "__main__":
    build_all()
