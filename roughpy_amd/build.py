"""Builds roughpy_amd/_lib/libroughpy_amd.so for gfx950 with hipcc.

Run as ``python -m roughpy_amd.build`` (or through __graft_entry__.build()).
Compiles every source under roughpy_amd/csrc in parallel and links one
shared library; the build runs without a GPU (hipcc cross-compiles).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_NAME = "libroughpy_amd.so"

ARCH = os.environ.get("ROUGHPY_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off (v_pk_fma_f32 and friends)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", "-I", os.path.join(HERE, "..", "include")]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _compile(src: str, obj: str, is_hip: bool) -> tuple:
    if is_hip:
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", src, "-o", obj]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", src, "-o", obj]
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src, p.returncode, p.stdout + p.stderr, time.time() - t0


def build(verbose: bool = False, jobs: int = 0) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(BUILD_DIR, exist_ok=True)
    hip, cpp = _sources()
    tasks = []
    for f in hip:
        tasks.append((os.path.join(CSRC, f), os.path.join(BUILD_DIR, f + ".o"), True))
    for f in cpp:
        tasks.append((os.path.join(CSRC, f), os.path.join(BUILD_DIR, f + ".o"), False))
    jobs = jobs or min(len(tasks), max(1, min(8, os.cpu_count() or 1)))
    errors = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, log, dt in ex.map(lambda t: _compile(*t), tasks):
            name = os.path.basename(src)
            if rc != 0:
                errors.append(f"--- {name} ---\n{log}")
            elif verbose:
                print(f"compiled {name} in {dt:.1f}s", flush=True)
                if log.strip():
                    print(log, flush=True)
    if errors:
        raise RuntimeError("roughpy_amd build failed:\n" + "\n".join(errors))
    objs = [t[1] for t in tasks]
    tmp = os.path.join(OUT_DIR, LIB_NAME + ".tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("roughpy_amd link failed:\n" + p.stdout + p.stderr)
    final = os.path.join(OUT_DIR, LIB_NAME)
    os.replace(tmp, final)
    return final


def clean() -> None:
    shutil.rmtree(BUILD_DIR, ignore_errors=True)


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
