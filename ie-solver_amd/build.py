"""Builds libiesolver_amd.so (HIP, gfx950) in-tree.

    python ie-solver_amd/build.py

Objects go to ie-solver_amd/_build/; the shared library lands next to this
file so that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libiesolver_amd.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
    # keep hipcc from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HIP_SOURCES = ["rs_kernels.hip", "rs_sweep.hip", "rs_holes.hip", "rs_eval.hip", "rs_matvec.hip", "rs_fmm.hip", "rs_mlp.hip", "capi.cpp"]
HOST_SOURCES = ["tree.cpp", "fmm_tree.cpp"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("command failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        cmd = ["hipcc"] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        if verbose and out:
            print(out)
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        _run(["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_cpp_tests()
    return LIB


def build_cpp_tests():
    """The C++ facade test program (tests/cpp), linked against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "ie_solver_gpu_test.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(root, "tests", "cpp", "ie_solver_gpu_test.bin")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"), src,
          "-o", out, "-L", HERE, "-liesolver_amd",
          "-Wl,-rpath,$ORIGIN/../../ie-solver_amd"])
    # CPU checks of the FMM tree and translations (host code only)
    fsrc = os.path.join(root, "tests", "cpp", "fmm_tree_test.cpp")
    if os.path.exists(fsrc):
        _run(["g++", "-O2", "-std=c++17", "-Wall", fsrc, os.path.join(CSRC, "fmm_tree.cpp"),
              "-o", os.path.join(root, "tests", "cpp", "fmm_tree_test.bin")])
    # CPU checks of the quadtree's local updates (host code only)
    tsrc = os.path.join(root, "tests", "cpp", "tree_update_test.cpp")
    if os.path.exists(tsrc):
        _run(["g++", "-O2", "-std=c++17", "-Wall", tsrc, os.path.join(CSRC, "tree.cpp"),
              "-o", os.path.join(root, "tests", "cpp", "tree_update_test.bin")])
    return out


if __name__ == "__main__":
    print(build(verbose=True))
