"""Build libffhip.so for gfx950 from csrc/ (in-tree, no JIT cache).

    python fourierflows.jl_amd/build.py [--jobs N] [--debug] [--sanitize]

Each .hip file is compiled to an object with hipcc in parallel, then linked
into ``fourierflows.jl_amd/lib/libffhip.so``.

``--sanitize`` builds ``lib/libffhip_san.so`` instead, with AddressSanitizer
and UndefinedBehaviorSanitizer on the host code only (plan factorisation,
twiddle generation, the ETDRK4 contour coefficients, argument checking); the
device code is built as usual.  It is for CPU runs of the host tests:

    FFHIP_LIB=fourierflows.jl_amd/lib/libffhip_san.so \
    LD_PRELOAD=$(python fourierflows.jl_amd/build.py --asan-runtime) \
    ASAN_OPTIONS=detect_leaks=0 python -m pytest tests/test_host.py
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libffhip.so")

ARCH = os.environ.get("FFHIP_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# host-only sanitizers: each -fsanitize applies to the host compilation only
SANITIZE_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]


def asan_runtime():
    """Path of the clang AddressSanitizer runtime that a sanitized build needs preloaded."""
    r = subprocess.run([hipcc(), "--print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    if r.returncode != 0 or not os.path.isabs(path):
        raise RuntimeError("clang AddressSanitizer runtime not found")
    return path


def common_flags(debug=False):
    f = [
        f"--offload-arch={ARCH}",
        "-std=c++17",
        "-fPIC",
        "-O1" if debug else "-O3",
        "-I", INCLUDE, "-I", CSRC,
        # packed FP32 instructions stay off for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        "-Wno-unknown-pragmas", "-Wno-unused-result",
    ]
    if debug:
        f.append("-g")
    return f


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src, flags, build_dir=BUILD):
    obj = os.path.join(build_dir, os.path.basename(src)[:-4] + ".o")
    cmd = [hipcc(), "-c", src, "-o", obj] + flags
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(
        ln for ln in (r.stderr or "").splitlines()
        if "warning" in ln and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln
    )
    return obj, warn


def build(jobs=None, debug=False, verbose=True, sanitize=False):
    build_dir = BUILD + ("_san" if sanitize else "")
    lib = os.path.join(LIBDIR, "libffhip_san.so") if sanitize else LIB
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    flags = common_flags(debug) + (SANITIZE_FLAGS if sanitize else [])
    srcs = sources()
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(compile_one, s, flags, build_dir): s for s in srcs}
        for f in cf.as_completed(futs):
            obj, warn = f.result()
            objs.append(obj)
            if warn and verbose:
                print(warn, file=sys.stderr)
    tmp = lib + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp, f"--offload-arch={ARCH}"] + sorted(objs)
    if sanitize:
        cmd += ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libsan"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--debug", action="store_true")
    ap.add_argument("--sanitize", action="store_true", help="host ASan + UBSan build: lib/libffhip_san.so")
    ap.add_argument("--asan-runtime", action="store_true", help="print the ASan runtime to preload and exit")
    a = ap.parse_args()
    if a.asan_runtime:
        print(asan_runtime())
    else:
        build(a.jobs, a.debug, sanitize=a.sanitize)
