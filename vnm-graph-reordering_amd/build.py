"""Build libvnm.so (HIP for gfx950 + host C++) and the drop-in `spmm` CLI.

    python vnm-graph-reordering_amd/build.py [--jobs N]

Objects go to vnm-graph-reordering_amd/_build/, the library to
vnm-graph-reordering_amd/lib/libvnm.so and the CLI to
vnm-graph-reordering_amd/bin/spmm.  hipcc cross-compiles gfx950 code objects
without a GPU present.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libvnm.so")
BIN = os.path.join(HERE, "bin", "spmm")
INC = os.path.join(HERE, "..", "include")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("VNM_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I" + INC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-munsafe-fp-atomics"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v3", "-I" + INC, "-Wall",
             "-Wno-unused-result", "-pthread"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    return p


def _compile(src):
    base = os.path.splitext(os.path.basename(src))[0]
    ext = os.path.splitext(src)[1]
    obj = os.path.join(OBJ, base + (".hip.o" if ext == ".hip" else ".o"))
    if ext == ".hip":
        cmd = [HIPCC, "-c", "-x", "hip", src, "-o", obj] + HIP_FLAGS
    else:
        cmd = ["g++", "-c", src, "-o", obj] + CXX_FLAGS + ["-I/opt/rocm/include"]
    _run(cmd)
    return obj


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    srcs = sorted(os.path.join(SRC, f) for f in os.listdir(SRC)
                  if f.endswith((".hip", ".cpp")) and f != "cli_main.cpp")
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "-pthread", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    cli = os.path.join(SRC, "cli_main.cpp")
    if os.path.exists(cli):
        _run(["g++", cli, "-o", BIN] + CXX_FLAGS +
             ["-L" + os.path.dirname(LIB), "-lvnm", "-Wl,-rpath,$ORIGIN/../lib"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
