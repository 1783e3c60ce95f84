"""Locating and loading the native libraries built in-tree.

``libsptc_gen.so`` is the host-only workload generator; ``libsptc_hip.so`` is
the product: hand-written gfx950 kernels behind the C ABI of
``include/sptc.h``.  Nothing here falls back to a CPU path: when the HIP
library is missing, :func:`hip_lib` raises.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
REPO_DIR = os.path.dirname(PKG_DIR)

GEN_LIB = os.path.join(LIB_DIR, "libsptc_gen.so")
HIP_LIB = os.path.join(LIB_DIR, "libsptc_hip.so")

_cache = {}


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path):
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run __graft_entry__.build() (or "
            f"`python -m gsparc_amd.build`) first")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[path] = lib
    return lib


def gen_lib():
    return _load(GEN_LIB)


def hip_lib():
    return _load(HIP_LIB)
