"""Build recipe for libbm_mpm.so (gfx950) and the C++ host test driver.

``python -m basements_amd.build`` compiles every HIP translation unit under
``basements_amd/csrc`` with hipcc for gfx950 and links them in-tree into
``basements_amd/lib/libbm_mpm.so``.  It always recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libbm_mpm.so")
HOSTTEST_SRC = os.path.join(ROOT, "tests", "cpp", "host_api_test.cpp")
HOSTTEST_BIN = os.path.join(LIBDIR, "host_api_test")
PLUGIN_SRC = os.path.join(PKG, "plugin", "bm_mpm_soil_plugin.c")
PLUGIN_LIB = os.path.join(LIBDIR, "libbm_mpm_soil_plugin.so")
# the reference's headers: the plugin form compiles against its plugin_capi.h
REF_INCLUDE = os.environ.get("BASEMENTS_REF_INCLUDE", "/root/reference/include")
PLUGIN_HDR = os.path.join("basements", "engine", "plugins", "plugin_capi.h")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Flags shared by every HIP TU.  Packed-FP32 instruction selection is turned
# off for the whole library; FP atomics use the hardware instructions
# (coarse-grained device memory only).
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

# (source, extra flags).  Reference-order kernels round every a*b+c separately.
UNITS = [
    ("mpm_ref_kernels.hip", ["-ffp-contract=off"]),
    ("mpm_fast_kernels.hip", []),
    ("mpm_split.hip", []),
    ("bm_mpm_api.hip", ["-ffp-contract=off"]),
    ("bm_batch.hip", []),
    ("bm_serve.hip", []),
    ("bm_diff.hip", []),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build_library(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmpdir = os.path.join(LIBDIR, "obj")
    os.makedirs(tmpdir, exist_ok=True)
    cmds, objs = [], []
    for src, extra in UNITS:
        obj = os.path.join(tmpdir, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *COMMON, *extra, "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src),
                     "-o", obj])
        objs.append(obj)
        if verbose:
            print(" ".join(cmds[-1]))
    # translation units compile independently: a few at a time
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "4")), os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    # roctx ranges (trace.hpp) resolve against the ROCm profiler SDK's roctx library
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-L", rocm_lib,
          "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{rocm_lib}"])
    os.replace(tmp, LIB)
    shutil.rmtree(tmpdir, ignore_errors=True)
    return LIB


def build_host_test(verbose: bool = False) -> str | None:
    """The C++ host-API test driver (links libbm_mpm.so through the C ABI only)."""
    if not os.path.exists(HOSTTEST_SRC):
        return None
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "host"),
           HOSTTEST_SRC, "-o", HOSTTEST_BIN, "-L", LIBDIR, "-lbm_mpm", f"-Wl,-rpath,{LIBDIR}",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return HOSTTEST_BIN


def plugin_cmd(out: str, ref_include: str = REF_INCLUDE, lang: str = "c") -> list:
    """Compile line of the reference-plugin form (plugin/bm_mpm_soil_plugin.c)
    against the reference's plugin header, linking libbm_mpm.so."""
    std = ["-x", "c", "-std=c99"] if lang == "c" else ["-x", "c++", "-std=c++17"]
    return ["gcc" if lang == "c" else "g++", *std, "-O2", "-Wall", "-Wextra", "-Werror", "-pedantic", "-shared",
            "-fPIC", "-isystem", ref_include, "-I", os.path.join(ROOT, "include"), PLUGIN_SRC, "-x", "none", "-L", LIBDIR,
            "-lbm_mpm", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN", "-o", out]


def build_plugin(verbose: bool = False) -> str | None:
    """libbm_mpm_soil_plugin.so, where the reference's plugin header is present."""
    if not os.path.exists(os.path.join(REF_INCLUDE, PLUGIN_HDR)):
        return None
    cmd = plugin_cmd(PLUGIN_LIB)
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return PLUGIN_LIB


def build(verbose: bool = False):
    lib = build_library(verbose)
    build_host_test(verbose)
    build_plugin(verbose)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
