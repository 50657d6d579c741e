"""Build recipe of libkylinv.so (HIP for gfx950) in this directory.

    python kylin-v_amd/build.py

Compiles the kernels and the C ABI with hipcc for gfx950 only, packed-FP32 VALU ops
disabled, and links rocSOLVER/rocBLAS for the FP64 bond SVD.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libkylinv.so")
CLI = os.path.join(HERE, "kln-tdvp")
SOURCES = ["kv_kernels.hip", "capi.cpp"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]
LIBS = ["-lrocsolver", "-lrocblas"]


def build(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp] + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libkylinv.so")
    os.replace(tmp, OUT)
    # the kln-tdvp command-line driver, linked against the library in this directory
    cli_tmp = CLI + ".tmp"
    cmd = [hipcc, "-O2", "-std=c++17", os.path.join(CSRC, "kln_tdvp.cpp"), "-o", cli_tmp,
           "-L" + HERE, "-lkylinv", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building kln-tdvp")
    os.replace(cli_tmp, CLI)
    return OUT


if __name__ == "__main__":
    build()
