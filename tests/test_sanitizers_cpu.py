"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

The reference builds its C++ layer with sanitizers on request
(CMakeLists.txt:81-86).  Here the library's host code (the fused-pass and
Pauli-word planners, the gate compiler, batched templates, argument checks)
and the oracle are built with ``-fsanitize=address,undefined`` on host code
only (``-Xarch_host``; device code is compiled normally and never runs) and
driven by ``tests/sanitize/host_driver.cpp``.  The build lands in
``pennylane-lightning-gpu_amd/_build/sanitize`` and is redone only when a
source changed.
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LGH = os.path.join(ROOT, "pennylane-lightning-gpu_amd", "csrc", "lgh")
OUT = os.path.join(ROOT, "pennylane-lightning-gpu_amd", "_build", "sanitize")
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang"
CLANGXX = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not all(os.path.exists(x) for x in (HIPCC, CLANG, CLANGXX)),
                                reason="ROCm compilers not installed")

SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer"]


def _stale(obj, srcs):
    return not os.path.exists(obj) or any(os.path.getmtime(s) > os.path.getmtime(obj) for s in srcs)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout[-4000:] + r.stderr[-4000:]


def build_driver():
    os.makedirs(OUT, exist_ok=True)
    headers = glob.glob(os.path.join(LGH, "*.hpp")) + [os.path.join(ROOT, "include", "lgh.h")]
    jobs, objs = [], []
    for src in sorted(glob.glob(os.path.join(LGH, "*.hip"))):
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        objs.append(obj)
        if _stale(obj, [src] + headers):
            jobs.append([HIPCC, "-O1", "-g", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                         "-Xclang", "-packed-fp32-ops", "-fPIC", *SAN_HOST, "-I", os.path.join(ROOT, "include"),
                         "-I", LGH, "-c", src, "-o", obj])
    drv = os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp")
    drv_obj = os.path.join(OUT, "host_driver.o")
    if _stale(drv_obj, [drv, os.path.join(ROOT, "include", "lgh.h")]):
        jobs.append([HIPCC, "-O1", "-g", "-std=c++17", "--offload-host-only", *SAN_HOST,
                     "-I", os.path.join(ROOT, "include"), "-c", drv, "-o", drv_obj])
    orc = os.path.join(ROOT, "oracle", "oracle.c")
    orc_obj = os.path.join(OUT, "oracle.o")
    if _stale(orc_obj, [orc, os.path.join(ROOT, "oracle", "oracle.h")]):
        jobs.append([CLANG, "-x", "c", "-std=gnu11", "-O1", "-g", "-fsanitize=address,undefined",
                     "-fno-omit-frame-pointer", "-Wno-unknown-pragmas", "-c", orc, "-o", orc_obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    exe = os.path.join(OUT, "host_driver")
    if jobs or not os.path.exists(exe):
        _run([CLANGXX, "-fsanitize=address,undefined", *objs, drv_obj, orc_obj, "-L/opt/rocm/lib", "-lamdhip64",
              "-lrocprofiler-sdk-roctx", "-lm", "-Wl,-rpath,/opt/rocm/lib", "-o", exe])
    return exe


@pytest.mark.timeout(900)
def test_host_code_under_asan_and_ubsan():
    exe = build_driver()
    env = dict(os.environ)
    # leaks of the HIP runtime's own one-time allocations are not ours
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert "runtime error" not in out, out[-4000:]
    assert r.returncode == 0 and "sanitized host checks passed" in out, out[-4000:]


def test_sanitize_flags_stay_on_host_code():
    """Every hipcc line of the sanitizer build puts each -fsanitize= behind
    -Xarch_host: GPU code is never built with a sanitizer."""
    cmd = [HIPCC, *SAN_HOST]
    for i, a in enumerate(cmd):
        if a.startswith("-fsanitize="):
            assert cmd[i - 1] == "-Xarch_host"
    assert shutil.which(HIPCC) or os.path.exists(HIPCC)
