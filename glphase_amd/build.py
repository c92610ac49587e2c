"""Builds the gfx950 shared library libglphase_amd.so in-tree with hipcc.

The library holds the HIP kernels, the extern "C" boundary declared in
include/glphase_amd.h and the C++ host classes.  Device and host code are
compiled with -ffp-contract=off: the forward recurrence must round every float
product and sum separately, as the reference CPU build does.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libglphase_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["glp_kernels.hip", "glp_work.hip", "glp_knn.hip", "glp_vq.hip", "glp_gen.hip",
           "glp_capi.cpp", "glp_host.cpp", "glp_io.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    # packed FP32 VALU ops stay off: they have given silently wrong results
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-pass-failed", "-I", os.path.join(ROOT, "include"),
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = [os.path.join(CSRC, "_obj_" + os.path.basename(s) + ".o") for s in srcs]
    cmds = [[HIPCC, "-x", "hip"] + [f for f in FLAGS if f != "-shared"] + ["-c", s, "-o", o]
            for s, o in zip(srcs, objs)]
    # the sources compile independently: at most 8 compilers at once
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"], verbose)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_cpp_tests(verbose=False):
    """The C++ drop-in test programs tests/cpp/test_hmmlike (the host
    classes) and tests/cpp/test_hmmlike_gsl (the reference-signature adapter
    of include/glphase_amd_hmmlike.hpp), linked to the library with an rpath
    to its directory."""
    outs = []
    for name in ("test_hmmlike", "test_hmmlike_gsl"):
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        out = os.path.join(ROOT, "tests", "cpp", name)
        cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
               src, "-o", out, "-L", HERE, "-lglphase_amd", "-Wl,-rpath," + HERE]
        _run(cmd, verbose)
        outs.append(out)
    return outs[0]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    if verbose and out.strip():
        print(out)


if __name__ == "__main__":
    print(build(verbose=True))
