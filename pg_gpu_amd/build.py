"""Build libpgld.so (HIP, gfx950) in-tree.

    python -m pg_gpu_amd.build           # or __graft_entry__.build()
    python -m pg_gpu_amd.build --asan    # + the AddressSanitizer build of the host code

Runs the code generator under csrc/ (gen_epilogue.py), compiles every csrc/*.hip to an
object with hipcc in parallel, and links pg_gpu_amd/libpgld.so.  The ASan build compiles
the host side instrumented (the device code is compiled as usual) into
_build/asan/libpgld_asan.so: it serves the CPU tests of the C ABI's argument checks and
the host-side band-schedule construction; load it with the clang ASan runtime preloaded.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpgld.so")
BUILD_DIR = os.path.join(os.path.dirname(HERE), "_build")
ARCH = os.environ.get("PGLD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 ops are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


ASAN_DIR = os.path.join(BUILD_DIR, "asan")
ASAN_OUT = os.path.join(ASAN_DIR, "libpgld_asan.so")
# host code only, instrumented; the sanitizer applies to the host pass
ASAN_FLAGS = ["-O1", "-g", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
              "-Xclang", "-packed-fp32-ops", "-fno-omit-frame-pointer", *"-Xarch_host -fsanitize=address".split()]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), "\n".join(lines)))
    return "\n".join(lines)


def build(verbose=False):
    for gen in ("gen_epilogue.py",):
        _run([sys.executable, os.path.join(CSRC, gen)])
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(BUILD_DIR, f[:-4] + ".o") for f in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
                for s, o in zip(srcs, objs)]
        for f in futs:
            out = f.result()
            if verbose and out:
                print(out)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, OUT)
    return OUT


def build_asan(verbose=False):
    for gen in ("gen_epilogue.py",):
        _run([sys.executable, os.path.join(CSRC, gen)])
    os.makedirs(ASAN_DIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(ASAN_DIR, f[:-4] + ".o") for f in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *ASAN_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
                for s, o in zip(srcs, objs)]
        for f in futs:
            out = f.result()
            if verbose and out:
                print(out)
    tmp = ASAN_OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", *"-Xarch_host -fsanitize=address".split(), *objs, "-o", tmp])
    os.replace(tmp, ASAN_OUT)
    return ASAN_OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    if "--asan" in sys.argv:
        print(build_asan(verbose="-v" in sys.argv))
