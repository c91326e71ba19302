"""Builds the HIP shared library libmoldyn_hip.so for gfx950, in-tree.

    python -m moleculardynamics_amd.build

Each .hip source is compiled with hipcc (one object per file, in parallel)
and linked into moleculardynamics_amd/lib/libmoldyn_hip.so. No cmake, no
JIT: the built library travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libmoldyn_hip.so")

SOURCES = ["md_capi.hip", "md_neigh.hip", "md_force.hip", "md_integrate.hip", "md_batch.hip",
           "md_batch_capi.hip", "md_water.hip", "md_water_capi.hip", "md_mix.hip",
           "md_mix_capi.hip", "md_noalloc.hip", "md_analysis.hip", "md_peaks.hip", "md_pme.hip"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instruction forms are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


CLI = os.path.join(HERE, "bin", "moldyn_run_gpu")


def build_cli(verbose: bool = True) -> str:
    """Compiles apps/moldyn_run_gpu.cpp (the reference driver's command line
    on the device path) against the library, into moleculardynamics_amd/bin."""
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    sha = "unknown"
    try:
        r = subprocess.run(["git", "-C", os.path.dirname(HERE), "rev-parse", "--short", "HEAD"],
                           capture_output=True, text=True, timeout=10)
        if r.returncode == 0 and r.stdout.strip():
            sha = r.stdout.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-DMOLDYN_GPU_GIT_SHA=\"{sha}\"", "-o", CLI,
           os.path.join(HERE, "apps", "moldyn_run_gpu.cpp"), "-L" + LIBDIR, "-lmoldyn_hip",
           "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"moldyn_run_gpu build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {CLI}")
    return CLI


SAN_DIR = os.path.join(OBJDIR, "sanitize")
SAN_EXE = os.path.join(SAN_DIR, "test_sanitize")
# host-only code with AddressSanitizer and UndefinedBehaviorSanitizer; each
# sanitizer flag applies to the host compilation only
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all"]


def build_sanitized(verbose: bool = True) -> str:
    """Builds tests/cpp/test_sanitize: the library with ASan + UBSan on its
    host code (device code as in the product build) and the CPU oracle with
    both sanitizers, linked into one driver that runs without a GPU."""
    os.makedirs(SAN_DIR, exist_ok=True)
    root = os.path.dirname(HERE)
    flags = FLAGS + ["-Xarch_host", "-g"] + SAN_HOST

    def one(src):
        obj = os.path.join(SAN_DIR, os.path.splitext(src)[0] + ".o")
        r = subprocess.run([HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"sanitized hipcc failed for {src}:\n{r.stderr}")
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(one, SOURCES))
    lib = os.path.join(SAN_DIR, "libmoldyn_hip_san.so")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *SAN_HOST, "-o", lib,
                        *objs], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitized link failed:\n{r.stderr}")
    clang = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib", "llvm", "bin", "clang++")
    if not os.path.exists(clang):
        clang = "/opt/rocm/lib/llvm/bin/clang++"
    osrc = [os.path.join(root, "oracle", f) for f in ("core.cpp", "lj.cpp", "integrate.cpp",
                                                       "ewald.cpp", "settle.cpp", "analysis.cpp")]
    cmd = [clang, "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-omit-frame-pointer", "-fno-sanitize-recover=all", "-o", SAN_EXE,
           os.path.join(root, "tests", "cpp", "test_sanitize.cpp"), *osrc, lib,
           "-Wl,-rpath," + SAN_DIR, "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"test_sanitize build failed:\n{r.stderr}")
    if verbose:
        print(f"built {SAN_EXE}")
    return SAN_EXE


def build_mirror_test(verbose: bool = True) -> str:
    """Compiles tests/cpp/test_mirror.cpp (the reference's test cases against
    the C++ drop-in classes of host/moldyn_gpu.hpp), linked to the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "test_mirror.cpp")
    outdir = os.path.join(root, "tests", "cpp", "_build")
    os.makedirs(outdir, exist_ok=True)
    exe = os.path.join(outdir, "test_mirror")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-o", exe, src, "-L" + LIBDIR, "-lmoldyn_hip",
           "-Wl,-rpath,$ORIGIN/../../../moleculardynamics_amd/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"test_mirror build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    build(verbose=True)
    build_mirror_test(verbose=True)
    build_cli(verbose=True)
    sys.exit(0)
