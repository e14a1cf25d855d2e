"""Build libhipotf.so (HIP, gfx950) in-tree: ``python -m binsim_amd.build``.

Every translation unit under csrc/ is compiled by hipcc for gfx950 in parallel into
``_build/`` and linked into ``binsim_amd/libhipotf.so``.  Packed FP32 instructions are
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).

``--asan`` builds ``binsim_amd/libhipotf_asan.so`` instead (objects in ``_build_asan/``): the host
code compiled with AddressSanitizer and UndefinedBehaviorSanitizer (each flag behind
``-Xarch_host``, so the device code is built as usual), for the CPU tests of the host paths.  It
needs the compiler's ASan runtime preloaded (``asan_runtime()``) and is selected with
``HIPOTF_LIBRARY``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libhipotf.so")
BUILD_ASAN = os.path.join(ROOT, "_build_asan")
LIB_ASAN = os.path.join(HERE, "libhipotf_asan.so")
SOURCES = ["basis.cpp", "ham.cpp", "capi.cpp", "sigma.hip", "rot.hip", "diag.hip", "davidson.cpp", "network.hip", "tables.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HIPOTF_ARCH", "gfx950")
FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result", "-I", CSRC]
# host-only sanitizers (every -fsanitize behind -Xarch_host: no sanitizer in GPU code)
ASAN_FLAGS = "-O1 -g -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer".split()


def asan_runtime() -> str | None:
    """The ASan runtime of the compiler that builds the library (to LD_PRELOAD), or None."""
    import glob
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def _compile(src: str, asan: bool = False) -> str:
    obj = os.path.join(BUILD_ASAN if asan else BUILD, os.path.splitext(src)[0] + ".o")
    flags = [f for f in FLAGS if f != "-O3"] + ASAN_FLAGS if asan else FLAGS
    cmd = [HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, asan: bool = False) -> str:
    out = LIB_ASAN if asan else LIB
    os.makedirs(BUILD_ASAN if asan else BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, asan), srcs))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    try:
        build(verbose=True, asan="--asan" in sys.argv[1:])
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
