"""Build libbz.so (all HIP sources under csrc/) for gfx950, in-tree.

    python -m boltzmann_amd.build        # or boltzmann_amd.build.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbz.so")
INCLUDE = os.path.abspath(os.path.join(HERE, "..", "include"))

SOURCES = ["bz_assembly.hip", "bz_solver0d.hip", "bz_solver0d_w8.hip", "bz_solver0d_w11.hip", "bz_solver0d_w12.hip",
           "bz_solver0d_w17.hip", "bz_solver0d_w20.hip", "bz_banded.hip", "bz_krylov.hip", "bz_ee.hip", "bz_noalloc.hip",
           "bz_plan0d.hip", "bz_vstep.hip", "bz_glow.hip", "bz_fvm.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # packed FP32 VALU ops are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


OBJDIR = os.path.join(HERE, "_build")


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel (one hipcc per source), then link."""
    from concurrent.futures import ThreadPoolExecutor
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    os.makedirs(OBJDIR, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = [hipcc()] + compile_flags + ["-I", INCLUDE, "-I", CSRC, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        return obj, res

    njobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(njobs) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = [res for _, res in results if res.returncode != 0]
    if failed:
        for res in failed:
            sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libbz.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + [o for o, _ in results]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libbz.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
