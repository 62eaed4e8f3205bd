"""Build recipe of libodx.so (HIP for gfx950) -- run as `python -m ode_amd.build`.

Compiles every source under ode_amd/csrc into one shared library in-tree
(ode_amd/libodx.so) with hipcc.  Packed-FP32 instruction selection is turned
off for all device code; FP contraction is off outside the network's kernels.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libodx.so")

SOURCES = ["odx_step.hip", "odx_space.hip", "odx_mlp.hip", "odx_batch.cpp", "odx_api.cpp", "odx_large.cpp"]
# Device code follows the reference's dSINGLE arithmetic without fused
# multiply-add contraction (an x86-64 build of ODE has none), so AABBs, poses
# and contacts are bit-exact; the SOR row updates re-enable contraction with
# a pragma.  The network's kernels keep the compiler's default.
NO_CONTRACT = {"odx_step.hip", "odx_space.hip", "odx_batch.cpp", "odx_api.cpp", "odx_large.cpp"}
EXAMPLES = ["frame_loop"]


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
        fp = ["-ffp-contract=off"] if os.path.basename(src) in NO_CONTRACT else []
        cmd = [hipcc, "-x", "hip", "-c", src, "-o", obj] + common + fp + list(extra)
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
    build_examples(verbose)
    return OUT


def build_examples(verbose=True):
    """Programs over the drop-in API (ode_amd/examples), linked against the
    in-tree libodx.so."""
    for name in EXAMPLES:
        src = os.path.join(HERE, "examples", name + ".cpp")
        out = os.path.join(HERE, "examples", name)
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"), src, "-o", out + ".tmp",
               "-L", HERE, "-lodx", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
