"""Builds the engine library and the ./ddm executable in-tree.

    python -m cuda_multiddm_amd.build        # library + CLI
Outputs (git-ignored, shipped with the tree):
    cuda_multiddm_amd/_build/libddm_amd.so
    cuda_multiddm_amd/_build/ddm
Host-only sources are compiled with g++ (-ffp-contract=off keeps the FP32
q-ring test rounding every step); HIP sources with hipcc for gfx950.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(OUT, "libddm_amd.so")
CLI = os.path.join(OUT, "ddm")

ARCH = os.environ.get("DDM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", "g++")

HOST_SRCS = ["ddm_host.cpp"]
HIP_SRCS = ["ddm_kernels.hip", "ddm_lag.hip", "ddm_engine.hip", "ddm_server.hip", "ddm_fit.hip", "ddm_fft.hip", "ddm_dft_mfma.hip"]
CLI_SRCS = ["ddm_cli.cpp", "ddm_run.cpp", "movie_reader.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: some of their forms have produced
    # wrong results on MI355X boxes without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{err}")
    return err


def _obj(src):
    return os.path.join(OUT, os.path.basename(src) + ".o")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    inc = [f"-I{INC}", f"-I{CSRC}"]
    jobs = []
    for s in HOST_SRCS:
        jobs.append([CXX, *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in HIP_SRCS:
        jobs.append([HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in CLI_SRCS:
        jobs.append([CXX, *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for msg in ex.map(_run, jobs):
            if verbose and msg.strip():
                print(msg, file=sys.stderr)
    objs = [_obj(s) for s in HOST_SRCS + HIP_SRCS]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    cli_objs = [_obj(s) for s in CLI_SRCS]
    _run([CXX, "-o", CLI + ".tmp", *cli_objs, f"-L{OUT}", "-lddm_amd", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
