"""Builds libflof_hip.so (the HIP kernels + C-ABI) for gfx950 with hipcc.

    python -m ofblend_amd.build          # or ofblend_amd.build.build()

Every source is compiled for gfx950 only, with -ffp-contract=off (no fused
multiply-add, so the FP32 arithmetic matches the CPU reference operation by
operation) and with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "flof_hip")
LIB = os.path.join(HERE, "libflof_hip.so")
SOURCES = ["flof_core.hip", "flof_solve.hip", "flof_frames.hip", "flof_prep.hip",
           "flof_warmnet.hip", "flof_post.hip", "flof_mesh.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    # hipcc's host pass does not know the GPU target feature: harmless
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False, jobs=None):
    """Compile every HIP source and link the shared library (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for s in SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC] + CFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o:
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
