"""Build recipes: HIP engine for gfx950 (hipcc), Fortran host (flang), CPU oracle (gcc)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_engine(verbose=False):
    out = os.path.join(HERE, "libfluxo_gpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(HERE, "csrc", "fx_engine.hip")], verbose)
    os.replace(tmp, out)
    return out


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


def build_all(verbose=False):
    build_engine(verbose)
    build_oracle(verbose)
