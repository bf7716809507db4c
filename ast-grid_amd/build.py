"""Builds the HIP/C++ library ast-grid_amd/astg/libastg.so for gfx950.

Every source is compiled with hipcc (device code for gfx950 only, packed FP32
instructions disabled), objects go to ast-grid_amd/_build/, and the shared
library is linked in-tree so that it travels with the repository.

    python ast-grid_amd/build.py [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "astg", "libastg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["octree.cpp", "scene.cpp", "weights.cpp", "project_host.cpp", "capi.cpp", "sweep_merge.cpp", "mic_plan.cpp", "frame_io.cpp",
           "consts.cpp", "reorder.cpp", "pcg.hip", "pcg_batch.hip", "plan_device.hip", "project.hip", "advect.hip",
           "linrec.hip", "reorder.hip", "ast_uniform.hip", "capi_ast.cpp"]
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I" + os.path.join(HERE, "..", "include"),
]


def compile_one(src, force=False):
    path = os.path.join(SRC, src)
    obj = os.path.join(OBJ, src + ".o")
    deps = [path] + [os.path.join(SRC, h) for h in os.listdir(SRC) if h.endswith(".h")]
    deps.append(os.path.join(HERE, "..", "include", "astg.h"))
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj, ""
    lang = ["-x", "hip"]
    cmd = [HIPCC] + lang + FLAGS + ["-c", path, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    return obj, p.stderr


def build(jobs=8, verbose=False, force=False):
    """force: compile every source even where its object is newer."""
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda src: compile_one(src, force), SOURCES))
    if verbose:
        for _, warn in results:
            if warn.strip():
                print(warn, file=sys.stderr)
    objs = [o for o, _ in results]
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", OUT + ".tmp"] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    os.replace(OUT + ".tmp", OUT)
    # C++ example of the CUDA-solver drop-in, linked against the library
    ex = os.path.join(HERE, "astg", "cuda_solver_shim")
    cmd = ["g++", "-std=c++17", "-O2", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "examples", "cuda_solver_shim.cpp"), "-o", ex + ".tmp",
           "-L" + os.path.dirname(OUT), "-lastg", "-Wl,-rpath,$ORIGIN"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("shim example build failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    os.replace(ex + ".tmp", ex)
    so = os.path.join(HERE, "astg", "libastg_shim_example.so")
    cmd = ["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-DASTG_SHIM_NO_MAIN",
           "-I" + os.path.join(HERE, "..", "include"), os.path.join(HERE, "examples", "cuda_solver_shim.cpp"),
           "-o", so + ".tmp", "-L" + os.path.dirname(OUT), "-lastg", "-Wl,-rpath,$ORIGIN"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("shim example build failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    os.replace(so + ".tmp", so)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
