"""Builds the native parts of the package in-tree.

1. ``liblgh.so``: the HIP kernels and the C ABI of ``include/lgh.h``, compiled
   with hipcc for gfx950 (one object per ``csrc/lgh/*.hip``, compiled in
   parallel).
2. ``lightning_gpu_qubit_ops``: the pybind11 module, compiled with g++ from
   ``csrc/bindings`` + ``csrc/host``; it includes only ``lgh.h`` and links
   ``liblgh.so`` through an ``$ORIGIN`` rpath.

Both land in ``pennylane_lightning_gpu/`` next to the Python sources.

Usage: ``python build_ext.py`` (or ``build()`` from Python).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
PKG = os.path.join(HERE, "pennylane_lightning_gpu")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("LGH_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", os.path.join(CSRC, "lgh"),
]
# The fused kernels branch on per-gate descriptor fields that are uniform
# across the wavefront; leaving those regions unstructurized keeps the
# register arrays in place across the branches instead of duplicating them.
FILE_FLAGS = {
    "lgh_fused.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions=true"],
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _newer(target, sources):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(s) <= t for s in sources)


def build_lgh(force=True, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "lgh", "*.hip")))
    headers = glob.glob(os.path.join(CSRC, "lgh", "*.hpp")) + [os.path.join(INCLUDE, "lgh.h")]
    objs = []
    jobs_list = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs_list.append([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(os.path.basename(s), []), "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, jobs_list))
    out = os.path.join(PKG, "liblgh.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-Wl,-soname,liblgh.so", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    return out


def module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "lightning_gpu_qubit_ops" + suffix)


def build_bindings():
    import pybind11

    src = os.path.join(CSRC, "bindings", "Bindings.cpp")
    out = module_path()
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", "-Wall",
        "-I", pybind11.get_include(), "-I", py_inc, "-I", INCLUDE, "-I", os.path.join(CSRC, "host"),
        src, "-o", tmp, "-L", PKG, "-llgh", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build(force=True):
    lib = build_lgh(force=force)
    mod = build_bindings()
    return lib, mod


if __name__ == "__main__":
    force = "--incremental" not in sys.argv
    for p in build(force=force):
        print("built", p)
