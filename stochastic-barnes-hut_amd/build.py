"""Build script of the package: compiles libsbh_hip.so (HIP, gfx950) and the
pybind11 plugin modules `sbh`, `sbh_gpu` and `sbh_gpu64` (the double build)
in place, next to this file.

    python stochastic-barnes-hut_amd/build.py

The plugin modules are host-only C++ (g++) linked against libsbh_hip.so
through its C ABI (include/sbh_hip.h) with an $ORIGIN rpath.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["tree_build.hip", "eval.hip", "eval_double.hip", "brute_mfma.hip", "serve.hip",
               "tree_io.hip", "capi.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared",
    # no packed-fp32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
LIB = os.path.join(PKG, "libsbh_hip.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_library() -> str:
    """Each source compiles to its own object (in parallel), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *compile_flags, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    jobs = max(1, min(len(HIP_SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def _ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX")


def build_module(name: str, source: str, link_lib: bool) -> str:
    import pybind11

    out = os.path.join(PKG, name + _ext_suffix())
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           os.path.join(CSRC, source), "-o", out + ".tmp"]
    if link_lib:
        cmd += [f"-L{PKG}", "-lsbh_hip", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all() -> list[str]:
    outs = [build_library()]
    outs.append(build_module("sbh", "py_sbh.cpp", link_lib=True))
    outs.append(build_module("sbh_gpu", "py_sbh_gpu.cpp", link_lib=True))
    outs.append(build_module("sbh_gpu64", "py_sbh_gpu64.cpp", link_lib=True))
    return outs


if __name__ == "__main__":
    for p in build_all():
        print(p)
