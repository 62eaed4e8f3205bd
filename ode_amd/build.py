"""Build recipe of libodx.so (HIP for gfx950) -- run as `python -m ode_amd.build`.

Compiles every source under ode_amd/csrc into one shared library in-tree
(ode_amd/libodx.so) with hipcc.  Packed-FP32 instruction selection is turned
off for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libodx.so")

SOURCES = ["odx_step.hip", "odx_mlp.hip", "odx_batch.cpp", "odx_api.cpp", "odx_large.cpp"]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=True, extra=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    common = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-I", os.path.join(HERE, "..", "include")]
    for src in sources():
        obj = os.path.join(HERE, "_build", os.path.basename(src) + ".o")
        cmd = [hipcc, "-x", "hip", "-c", src, "-o", obj] + common + list(extra)
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
