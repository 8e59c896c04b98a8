"""Builds the gfx950 plugin library (libwinoconv.so) and the C++ harness in-tree with hipcc.

Usage: python 2024pra_conv_amd/build_lib.py   (or __graft_entry__.build()).
Compiles every time it is called.

       python 2024pra_conv_amd/build_lib.py --sanitize
builds a second copy of the library, lib/sanitize/libwinoconv.so, whose host code (plugin
dispatch, tuned-table and U-cache file parsing) carries AddressSanitizer and
UndefinedBehaviorSanitizer checks (device code is built as usual). The CPU tests run against it
with the sanitizer runtime preloaded:
       LD_PRELOAD=$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so) \
       ASAN_OPTIONS=detect_leaks=0 WINOCONV_LIB=2024pra_conv_amd/lib/sanitize/libwinoconv.so \
       python -m pytest tests -m "not gpu"
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "lib", "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

LIB_SOURCES = [
    "wino_transforms.hip",
    "wino_gemm.hip",
    "wino_fp8.hip",
    "wino_fused.hip",
    "direct_pack.hip",
    "direct_f64.hip",
    "conv2d_plugin.hip",
    "tuner.hip",
    "batch.hip",
    "u_cache.hip",
    "mlp.hip",
    "static_scratch.hip",
    "wino43.hip",
]

# Packed FP32 VALU instructions are kept out of all device code: the compiler would otherwise
# form them from float2-shaped arithmetic, and they are not needed for throughput here.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed: {cmd[-1]}")
    # the host pass does not know the device-only feature flag; that warning is harmless
    noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


# host-only sanitizers: each -fsanitize= applies to the host pass alone
SANITIZE_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


def _compile(src: str, objdir: str = OBJDIR, extra: tuple = ()) -> str:
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *COMMON, *DEVICE_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_sanitized(verbose: bool = False) -> str:
    """The library with host-side ASan + UBSan, in lib/sanitize/ (the product library is untouched)."""
    objdir = os.path.join(OBJDIR, "sanitize")
    outdir = os.path.join(LIBDIR, "sanitize")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(outdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(LIB_SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, objdir, tuple(SANITIZE_HOST)), LIB_SOURCES))
    lib = os.path.join(outdir, "libwinoconv.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fsanitize=address", "-fsanitize=undefined",
          "-shared-libasan", *objs, "-o", lib])
    if verbose:
        print(f"built {lib}")
    return lib


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(LIB_SOURCES))) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    lib = os.path.join(LIBDIR, "libwinoconv.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, lib)
    # C++ harness (counterpart of the reference's main.cpp), linked against the library
    demo = os.path.join(LIBDIR, "conv2dfp16demo")
    _run([HIPCC, *COMMON, *DEVICE_FLAGS, os.path.join(CSRC, "demo_main.cpp"),
          f"-L{LIBDIR}", "-lwinoconv", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN",
          "-o", demo])
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        build_sanitized(verbose=True)
    else:
        build(verbose=True)
