"""Build the gfx950 shared library ``libwapfdtd.so`` in-tree with hipcc.

``python -m wideanglepropagation_amd.build`` (or ``build()``) compiles
``csrc/*.hip`` for gfx950 only.  The library is loaded with ctypes by
``wideanglepropagation_amd._lib``.
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libwapfdtd.so")
CLI_PATH = os.path.join(PKG_DIR, "fdtd_3d_bin")
SOURCES = ["fdtd3d_capi.hip"]
CLI_SOURCE = "fdtd_3d_main.cpp"

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # keep packed-FP32 VALU forms out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(PKG_DIR, "..", "include"),
                                     "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc build of libwapfdtd.so failed")
    os.replace(tmp, LIB_PATH)
    build_cli(verbose)
    return LIB_PATH


def build_cli(verbose=False):
    """Host-only ``fdtd_3d_bin`` (the reference's argv/file contract) linked to the library."""
    tmp = CLI_PATH + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(PKG_DIR, "..", "include"),
           "-o", tmp, os.path.join(CSRC, CLI_SOURCE), "-L", PKG_DIR, "-lwapfdtd",
           "-Wl,-rpath," + PKG_DIR]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build of fdtd_3d_bin failed")
    os.replace(tmp, CLI_PATH)
    return CLI_PATH


if __name__ == "__main__":
    print(build(verbose=True))
