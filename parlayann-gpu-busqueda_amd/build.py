"""Build the native libraries in-tree.

    libpann.so        HIP (gfx950) kernels + the C ABI of include/pann.h
    libpann_tools.so  host C++ tools: synthetic data, graph builders (include/pann_tools.h)

Usage: python build.py [--jobs N]
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
REPO = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

# Packed-FP32 ops are switched off for all device code.
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(REPO, 'include')}",
]
HOST_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-march=x86-64-v3", "-pthread", "-Wall",
              f"-I{os.path.join(REPO, 'include')}"]

PANN_HIP_SOURCES = ["beam_flat.hip", "beam_hnsw.hip", "bf_knn.hip", "rows.hip", "build_gpu.hip", "build_gpu_trees.hip", "pann_api.cpp", "pann_batch.cpp", "pann_server.cpp", "pann_noalloc.cpp", "formats.cpp"]
TOOLS_SOURCES = ["tools.cpp", "build_vamana.cpp", "build_hnsw.cpp", "build_hcnng.cpp",
                 "build_pynndescent.cpp", "formats.cpp"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit(f"build failed: {os.path.basename(cmd[-1])}")
    return r


# Per-file device flags: the brute-force GEMM keeps MFMA results in VGPRs (no AGPR reads
# before each compare).
FILE_FLAGS = {"bf_knn.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def compile_one(src, obj, hip, extra=()):
    extra = list(extra) + (FILE_FLAGS.get(os.path.basename(src), []) if hip else [])
    if hip:
        cmd = [HIPCC] + HIP_FLAGS + list(extra) + ["-c", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + HIP_FLAGS + list(extra) + ["-x", "hip", "-c", src, "-o", obj]
    else:
        cmd = [CXX] + HOST_FLAGS + ["-c", src, "-o", obj]
    run(cmd)
    return obj


def build_profile(jobs=8):
    """libpann_prof.so: the same library with per-section cycle counters compiled in."""
    os.makedirs(OBJ, exist_ok=True)
    tasks = [(os.path.join(CSRC, s), os.path.join(OBJ, "prof_" + s + ".o"), True, ("-DPANN_PROFILE",))
             for s in PANN_HIP_SOURCES]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda t: compile_one(*t), tasks))
    out = os.path.join(HERE, "libpann_prof.so")
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)


def build(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    tasks = []
    for s in PANN_HIP_SOURCES:
        tasks.append((os.path.join(CSRC, s), os.path.join(OBJ, "hip_" + s + ".o"), True))
    tools = [s for s in TOOLS_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    for s in tools:
        tasks.append((os.path.join(CSRC, s), os.path.join(OBJ, "host_" + s + ".o"), False))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda t: compile_one(*t), tasks))
    hip_objs = [o for o, t in zip(objs, tasks) if t[2]]
    host_objs = [o for o, t in zip(objs, tasks) if not t[2]]
    out = os.path.join(HERE, "libpann.so")
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp"] + hip_objs)
    os.replace(out + ".tmp", out)
    if host_objs:
        out2 = os.path.join(HERE, "libpann_tools.so")
        run([CXX, "-shared", "-fPIC", "-pthread", "-o", out2 + ".tmp"] + host_objs)
        os.replace(out2 + ".tmp", out2)


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("--profile", action="store_true", help="also build libpann_prof.so")
    a = ap.parse_args()
    if a.clean:
        clean()
    build(a.jobs)
    if a.profile:
        build_profile(a.jobs)
    print("built", os.path.join(HERE, "libpann.so"))
