"""Build and load the native libraries (in-tree, so the built .so files travel
with the source tree).

  libflsg.so          HIP kernels + C ABI (include/flsg.h), hipcc for gfx950
  libflsg_graphgen.so host-only synthetic graph generators (gcc)

``build()`` always recompiles; ``load_*()`` loads what is built and compiles
only when nothing is.  The product path never falls back to anything else:
if libflsg.so cannot be loaded the call raises.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_DIR, "csrc")
REPO = os.path.dirname(PKG_DIR)
INCLUDE = os.path.join(REPO, "include")
LIB_PATH = os.path.join(PKG_DIR, "libflsg.so")
GRAPHGEN_PATH = os.path.join(PKG_DIR, "libflsg_graphgen.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FLSG_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit parity with the reference's FP64 arithmetic: never fuse mul+add
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]

_lib = None
_graphgen = None


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout[-4000:] + r.stderr[-8000:])
    return r


def build_graphgen() -> str:
    src = os.path.join(CSRC, "graphgen.c")
    tmp = GRAPHGEN_PATH + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-o", tmp, src, "-lm"])
    os.replace(tmp, GRAPHGEN_PATH)
    return GRAPHGEN_PATH


def build_lib() -> str:
    """One object per source, compiled in parallel, then linked (each
    translation unit's device code is self-contained: no relocatable device
    code)."""
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(CSRC, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"]

    def obj(src):
        o = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        _run([HIPCC, *cflags, f"-I{INCLUDE}", "-c", "-o", o, src])
        return o

    workers = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(obj, srcs))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_all() -> None:
    build_graphgen()
    build_lib()


def load_graphgen():
    global _graphgen
    if _graphgen is None:
        if not os.path.exists(GRAPHGEN_PATH):
            build_graphgen()
        _graphgen = ctypes.CDLL(GRAPHGEN_PATH)
        _graphgen.flsg_gen_chung_lu.restype = ctypes.c_int
        _graphgen.flsg_gen_pref_attach.restype = ctypes.c_int
        _graphgen.flsg_spanning_forest.restype = ctypes.c_int
    return _graphgen


def load_lib():
    """ctypes handle of libflsg.so (raises if it cannot be built or loaded)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build_lib()
        from . import _abi
        _lib = _abi.bind(ctypes.CDLL(LIB_PATH))
    return _lib
