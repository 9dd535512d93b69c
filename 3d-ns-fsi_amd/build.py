"""Builds libnsfsi.so (HIP kernels for gfx950 + host library + C-ABI) in-tree.

Usage: python 3d-ns-fsi_amd/build.py [--jobs N]
Every source is compiled by hipcc for gfx950 only. Files whose arithmetic must match a serial FP64
evaluation bit for bit (classification, ghost geometry, boundary conditions, host set-up) are built
with -ffp-contract=off.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnsfsi.so")

SOURCES = [
    ("ConfigSettings.cpp", True),
    ("capi.cpp", True),
    ("mesh.hip", True),
    ("solver.hip", True),
    ("stage_plain.hip", True),
    ("stage_fused.hip", False),
    ("reduce.hip", True),
    ("batch.cpp", False),
    ("adjoint.hip", True),
    ("adjoint_host.cpp", True),
    ("noalloc.hip", True),
]

ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (hand-off to the compiler only emits scalar/vector FP32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_one(src, strict):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + (["-ffp-contract=off"] if strict else []) + ["-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out, r.stderr


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [(s, strict) for s, strict in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out, err in ex.map(lambda a: compile_one(*a), srcs):
            objs.append(out)
            if verbose and err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host_programs()
    return LIB


# Host programs over the C-ABI, compiled with g++ against libnsfsi.so (found next to them, or through the
# rpath): the command-line program of the reference's main, and the drop-in test program of the test suite.
ROOT = os.path.dirname(HERE)
HOST_PROGRAMS = [
    (os.path.join(HERE, "host", "main.cpp"), os.path.join(HERE, "nsfsi_run")),
    (os.path.join(ROOT, "tests", "cpp", "solver_dropin.cpp"), os.path.join(ROOT, "tests", "cpp", "solver_dropin")),
]


def build_host_programs():
    for src, out in HOST_PROGRAMS:
        if not os.path.exists(src):
            continue
        rpath = os.path.relpath(HERE, os.path.dirname(out))
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", out + ".tmp",
               "-L", HERE, "-lnsfsi", f"-Wl,-rpath,$ORIGIN/{rpath}", "-Wl,-rpath-link,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host program build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
