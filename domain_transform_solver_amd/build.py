"""Build recipe for libdts_amd.so (gfx950 HIP kernels + C ABI + C++ class API).

The library is built in-tree next to this file with hipcc; nothing is JIT
compiled at run time.  ``python -m domain_transform_solver_amd.build`` builds it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
LIB_PATH = os.path.join(PKG_DIR, "libdts_amd.so")

HIP_SOURCES = ["dts.hip", "hfbs.hip", "batch.hip", "adjoint.hip", "rf.hip"]
CXX_SOURCES = ["cxx_api.cpp"]

# Packed-FP32 instructions are disabled for all device code; -O3 hipcc would
# otherwise form v_pk_* FP32 ops from ordinary float code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every source into libdts_amd.so (always recompiles)."""
    objs = []
    obj_dir = os.path.join(PKG_DIR, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()
    srcs = [s for s in HIP_SOURCES + CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    for src in srcs:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [cc, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c",
                   os.path.join(CSRC, src), "-o", obj]
        else:
            cmd = [cc, "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE, "-I", CSRC,
                   "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    # roctx (named ranges around the C ABI calls) from the ROCm install.
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(cc))), "lib")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L" + rocm_lib, "-Wl,-rpath," + rocm_lib, "-lrocprofiler-sdk-roctx"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB_PATH)
    build_dropin(verbose)
    build_cli(verbose)
    return LIB_PATH


DROPIN_SRC = os.path.join(REPO_DIR, "tests", "cxx", "dropin_solve.cpp")
DROPIN_EXE = os.path.join(REPO_DIR, "tests", "cxx", "dropin_solve.exe")
REFUSE_SRC = os.path.join(REPO_DIR, "tests", "cxx", "reference_usage.cpp")
REFUSE_EXE = os.path.join(REPO_DIR, "tests", "cxx", "reference_usage.exe")


def build_dropin(verbose: bool = False) -> str | None:
    """The C++ drop-in programs: reference-style class API from include/ only,
    linked against libdts_amd.so with a relative rpath."""
    if not os.path.exists(DROPIN_SRC):
        return None
    for src, exe in ((DROPIN_SRC, DROPIN_EXE), (REFUSE_SRC, REFUSE_EXE)):
        if not os.path.exists(src):
            continue
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unused-parameter", "-I", INCLUDE, src,
               "-o", exe, "-L", PKG_DIR, "-ldts_amd",
               "-Wl,-rpath,$ORIGIN/../../domain_transform_solver_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    return DROPIN_EXE


CLI_DIR = os.path.join(PKG_DIR, "cli")
BIN_DIR = os.path.join(PKG_DIR, "bin")
CLI_PROGRAMS = ["depth_super_res_cli", "colorization_cli", "stereo_cli"]


def build_cli(verbose: bool = False) -> list[str]:
    """The command-line programs of the reference (depth super-resolution,
    colorization, stereo) on the C++ class API: g++ against include/ and
    cli/, linked to libdts_amd.so (relative rpath) and zlib."""
    os.makedirs(BIN_DIR, exist_ok=True)
    io_src = os.path.join(CLI_DIR, "image_io.cpp")
    out = []
    for name in CLI_PROGRAMS:
        src = os.path.join(CLI_DIR, name + ".cpp")
        if not os.path.exists(src):
            continue
        exe = os.path.join(BIN_DIR, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", CLI_DIR, src, io_src,
               "-o", exe, "-L", PKG_DIR, "-ldts_amd", "-lz", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        out.append(exe)
    out.append(build_codec_harness(verbose))
    return out


HARNESS_SRC = os.path.join(REPO_DIR, "tests", "cxx", "png_dump.cpp")
HARNESS_EXE = os.path.join(REPO_DIR, "tests", "cxx", "png_dump.exe")
HARNESS_SAN_EXE = os.path.join(REPO_DIR, "tests", "cxx", "png_dump_san.exe")


def build_codec_harness(verbose: bool = False, sanitize: bool = False) -> str:
    """The codec test harness (tests/cxx/png_dump.cpp + cli/image_io.cpp), CPU
    only; with sanitize, an AddressSanitizer + UndefinedBehaviorSanitizer
    build that aborts on the first finding."""
    io_src = os.path.join(CLI_DIR, "image_io.cpp")
    exe = HARNESS_SAN_EXE if sanitize else HARNESS_EXE
    flags = (["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
              "-fno-sanitize-recover=all"] if sanitize else ["-O2"])
    cmd = ["g++", *flags, "-std=c++17", "-Wall", "-I", CLI_DIR, HARNESS_SRC, io_src, "-o", exe,
           "-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return exe


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB_PATH)
