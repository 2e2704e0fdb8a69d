"""Build the HIP library ``forcepho_amd/libfpx.so`` for gfx950.

The library is compiled in-tree with ``hipcc`` (no JIT cache), so a built
tree can be copied to a GPU host and used without a compiler.
"""

import os
import shutil
import subprocess

__all__ = ["PKG_DIR", "REPO_DIR", "LIB_PATH", "build_library"]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
INCLUDE_DIR = os.path.join(REPO_DIR, "include")
CSRC_DIR = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libfpx.so")

SOURCES = ["fpx_kernel.hip", "fpx_records.hip", "fpx_evaluator.hip", "fpx_static.hip"]
HEADERS = ["fpx_gauss.h", "fpx_internal.h"]

# Packed FP32 (v_pk_*_f32) is disabled for all device code: the compiler
# otherwise forms those instructions from scalar float code by itself.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build_library(verbose=False):
    """Compile libfpx.so; always recompiles.  Returns the library path."""
    srcs = [os.path.join(CSRC_DIR, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [_hipcc(), *HIPCC_FLAGS, "-I" + INCLUDE_DIR, "-I" + CSRC_DIR, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns about the device-only target feature; drop that noise
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
