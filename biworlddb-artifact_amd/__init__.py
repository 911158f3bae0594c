"""bwgpu -- an MI355X-native GPU batch path for BiWorldDB-style version-tree range sums.

The product is native code built in-tree (see Makefile):
  build/libbwgpu.so   HIP kernels for gfx950 behind a C-ABI (reference drop-in symbols of
                      gpu_v7.h / gpu_v7b.h / gpu_session.h, plus the bwt_* tree-path entry points)
  build/libbwhost.so  host plan / workload code
  build/bw_tree       timed worker of the tree path (configs 1-5)
  build/bw_panel      timed worker of the reference drop-in path (reference-native workloads)
This module only locates those artefacts, builds them, and loads the C-ABI with ctypes.  There
is no Python or CPU fallback for any entry point: a missing library raises.
"""
import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
BUILD_DIR = os.path.join(PKG_DIR, "build")
LIB_GPU = os.path.join(BUILD_DIR, "libbwgpu.so")
LIB_HOST = os.path.join(BUILD_DIR, "libbwhost.so")
BW_TREE = os.path.join(BUILD_DIR, "bw_tree")
BW_PANEL = os.path.join(BUILD_DIR, "bw_panel")
BW_BATCH = os.path.join(BUILD_DIR, "bw_batch")
BW_TUNE = os.path.join(BUILD_DIR, "bw_tune")
BW_SERVE = os.path.join(BUILD_DIR, "bw_serve")
TUNED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
INCLUDE_DIR = os.path.join(REPO_DIR, "include")


def build(jobs=8):
    """Compile every HIP source for gfx950 and the host code (always recompiles)."""
    subprocess.run(["make", "-C", PKG_DIR, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", PKG_DIR, f"-j{jobs}"], check=True)


def require_built():
    missing = [p for p in (LIB_GPU, LIB_HOST, BW_TREE, BW_PANEL, BW_BATCH, BW_TUNE, BW_SERVE) if not os.path.exists(p)]
    if missing:
        raise RuntimeError("bwgpu native build missing (run make -C biworlddb-artifact_amd): " + ", ".join(missing))


def load_gpu_lib():
    """ctypes handle on the C-ABI library (raises if it is not built)."""
    if not os.path.exists(LIB_GPU):
        raise RuntimeError(f"bwgpu: {LIB_GPU} is not built")
    return ctypes.CDLL(LIB_GPU)


def declared_symbols():
    """Every function name declared in include/*.h (the exported C-ABI)."""
    import re

    names = []
    for fn in sorted(os.listdir(INCLUDE_DIR)):
        if not fn.endswith(".h"):
            continue
        with open(os.path.join(INCLUDE_DIR, fn)) as f:
            txt = f.read()
        names += re.findall(r"^\s*(?:extern \"C\" )?int\s+(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))
