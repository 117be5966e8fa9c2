"""Builds the radik_amd C-ABI library for gfx950 with hipcc, in-tree.

Output: radik_amd/lib/libradik_amd.so (git-ignored; it travels with the tree).
Object files go to radik_amd/_build/.  Run ``python -m radik_amd.build``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libradik_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = [
    "rk_kernels.hip",
    "rk_api.cpp",
    "rk_gen.hip",
    "rk_tune.cpp",
    "rk_serve.cpp",
    "rk_knn.hip",
    "rk_io.cpp",
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-I" + os.path.join(HERE, "..", "include"),
    f"--offload-arch={ARCH}",
    # packed fp32 instructions are kept out of every HIP code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str, build_dir: str = BUILD, defines: tuple = ()) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *defines, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        # host-pass warnings about the packed-fp32 feature are expected
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose: bool = True, out: str | None = None, defines: tuple = ()) -> str:
    """Compiles every source and links the library (to ``out`` if given,
    e.g. a tuning variant built with extra ``defines``)."""
    lib = out or LIB
    build_dir = BUILD if not defines else BUILD + "_" + "_".join(d.strip("-D") for d in defines)
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, defines), SOURCES))
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    build(out=a.out, defines=tuple("-D" + d for d in a.defines))
