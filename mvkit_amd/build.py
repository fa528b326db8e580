"""Builds the MI355X NCC library in-tree: mvkit_amd/libmvkit_amd.so.

HIP sources are compiled for gfx950 with hipcc; host C++ with g++ against
the HIP runtime headers.  Both use -ffp-contract=off so FP64 expressions are
evaluated in the reference's operation order (no fused multiply-adds).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmvkit_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]
CXX_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
             "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _compile(src: str) -> str:
    obj = os.path.join(OUT, src + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-c", path, "-o", obj]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    hip, cpp = _sources()
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, hip + cpp))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_program(source: str, output: str) -> str:
    """Compiles a C++ program that uses the library through include/mv_ncc.hpp
    (host code only: g++, linked against libmvkit_amd.so with an rpath to
    the package directory, so the program runs from where it was built)."""
    include = os.path.join(os.path.dirname(HERE), "include")
    rpath = os.path.relpath(HERE, os.path.dirname(os.path.abspath(output)))
    tmp = output + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-pedantic", "-Werror", f"-I{include}", source,
           "-o", tmp, f"-L{HERE}", "-lmvkit_amd", f"-Wl,-rpath,$ORIGIN/{rpath}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, output)
    return output


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
