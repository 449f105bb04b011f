"""Builds libreffine_hip.so in-tree (reffine_amd/_build/).

Host sources are compiled with hipcc as C++ (no device code); the
ahead-of-time kernels in csrc/kernels/rf_aot.hip are compiled for gfx950.
The device skeleton headers are embedded as string literals for hiprtc,
which specialises them per query at run time.

Usage: python -m reffine_amd.build   (or reffine_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libreffine_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["ir.cpp", "lower.cpp", "hipgen.cpp", "engine.cpp", "engine_batch.cpp",
                "arrow_device.cpp", "capi.cpp"]
DEVICE_SOURCES = ["kernels/rf_aot.hip", "kernels/rf_group.hip", "noalloc.hip"]
EMBED = {"kRfAbiSrc": "kernels/rf_abi.h", "kRfDeviceSrc": "kernels/rf_device.h"}

COMMON = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"-I{CSRC}", f"-I{BUILD}", f"-I{os.path.join(HERE, '..', 'include')}",
          f"-I{os.path.join(ROCM, 'include')}"]
# gfx950 device flags: packed FP32 instructions disabled, no FMA contraction
# so FP64 arithmetic rounds exactly as the C text says.
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-ffp-contract=off", "-munsafe-fp-atomics",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _embed() -> None:
    os.makedirs(BUILD, exist_ok=True)
    parts = []
    for name, rel in EMBED.items():
        with open(os.path.join(CSRC, rel)) as f:
            text = f.read()
        if ")RFSRC\"" in text:
            raise RuntimeError(f"{rel} contains the raw-string delimiter")
        parts.append(f'static const char* {name} = R"RFSRC({text})RFSRC";\n')
    out = os.path.join(BUILD, "rf_embed.inc")
    data = "".join(parts)
    if not os.path.exists(out) or open(out).read() != data:
        with open(out, "w") as f:
            f.write(data)


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def build(verbose: bool = False) -> str:
    """Compiles every source (always) and links the shared library."""
    _embed()
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append([HIPCC, "-x", "c++", "-D__HIP_PLATFORM_AMD__", *COMMON, "-c",
                     os.path.join(CSRC, src), "-o", obj])
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append([HIPCC, "-x", "hip", *COMMON, *DEVICE_FLAGS, "-c",
                     os.path.join(CSRC, src), "-o", obj])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as ex:
        for j in jobs:
            if verbose:
                print(" ".join(j), file=sys.stderr)
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
          f"-L{os.path.join(ROCM, 'lib')}", "-lhiprtc", "-lamdhip64",
          f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"])
    os.replace(tmp, LIB)
    return LIB


C_ABI_TEST = os.path.join(HERE, "..", "tests", "c_abi")


def build_c_abi_test() -> str:
    """Compiles the C host program of tests/c_abi against the library (gcc,
    no Python, no HIP headers: only include/reffine_hip.h)."""
    out_dir = os.path.join(C_ABI_TEST, "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "rf_c_scan")
    _run(["gcc", "-O2", "-std=c11", "-Wall", f"-I{os.path.join(HERE, '..', 'include')}",
          os.path.join(C_ABI_TEST, "rf_c_scan.c"), "-o", exe, f"-L{BUILD}", "-lreffine_hip",
          f"-Wl,-rpath,{BUILD}", "-lm"])
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
