"""Builds the gfx950 shared library ``chidg_amd/libchidg_hip.so`` in-tree.

Every translation unit is compiled by hipcc for gfx950 only; packed-FP32 ops
are disabled for device code.  Objects go to ``chidg_amd/_build/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchidg_hip.so")
SOURCES = ["cdg_api.hip", "k_stream.hip", "k_dfapply.hip", "k_rsapply.hip", "k_factor.hip", "k_krylov.hip",
           "k_synth.hip", "k_multi.hip", "k_euler.hip", "cdg_batch.hip", "cdg_serve.hip", "cdg_noalloc.hip", "euler_jac.cpp"]
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
         "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):   # host-only C++ (no device code)
        rocm = os.environ.get("ROCM_PATH", os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))))
        cmd = [HIPCC, "-x", "c++", "-O3", "-fPIC", "-std=c++17", "-Wall", "-pthread", "-D__HIP_PLATFORM_AMD__",
               "-I", os.path.join(rocm, "include"), "-I", os.path.join(HERE, "..", "include"),
               "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    noise = "unknown target feature"
    msg = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l and "packed-fp32" not in l)
    if msg.strip():
        print(msg, file=sys.stderr)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    # every symbol of this library must resolve inside it (or in the HIP / C++ runtimes)
    r = subprocess.run(["nm", "-D", "--undefined-only", tmp], capture_output=True, text=True)
    own = [l.split()[-1] for l in r.stdout.splitlines() if l.split() and "@" not in l.split()[-1]
           and not l.split()[-1].startswith(("hip", "__hip", "_ITM", "__gmon", "__cxa", "_Unwind", "__amd"))]
    if own:
        os.remove(tmp)
        raise RuntimeError(f"unresolved symbols in {LIB}: {own}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
