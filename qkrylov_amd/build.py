"""Build the engine shared library qkrylov_amd/libqkrylov_amd.so with hipcc
for gfx950 (one object per source, compiled in parallel, then linked).

    python qkrylov_amd/build.py [--jobs N] [--debug]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqkrylov_amd.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = [
    "qk_kernels.hip",
    "qk_model.cpp",
    "qk_engine.cpp",
    "qk_api.cpp",
    "qk_c_api.cpp",
    "qk_serving.cpp",
    "qk_eigen.cpp",
    "qk_split.cpp",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags(debug=False):
    f = [
        "--offload-arch=gfx950",
        "-std=c++20",
        "-fPIC",
        "-O1" if debug else "-O3",
        "-g" if debug else "-DNDEBUG",
        # no packed FP32 instructions: some of their forms have given wrong
        # results on these GPUs without any error
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I" + os.path.join(ROOT, "include"),
        "-I" + CSRC,
        "-Wno-unused-command-line-argument",
    ]
    return f


def compile_one(src, debug):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [hipcc()] + flags(debug) + ["-c", os.path.join(CSRC, src), "-o", obj]
    if not src.endswith(".hip"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=None, debug=False, verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, debug), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args()
    try:
        build(a.jobs, a.debug)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
