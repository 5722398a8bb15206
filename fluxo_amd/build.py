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


# per-degree kernel units (fx_tu.hip): degrees 2..5 x kernel groups 0..3
TU_DEGREES = (2, 3, 4, 5)
TU_GROUPS = (0, 1, 2, 3)


def build_engine(verbose=False, jobs=None):
    """libfluxo_gpu.so from fx_engine.hip (context, C ABI, server, non-template kernels) and
    one object per (degree, kernel group) of fx_tu.hip, compiled in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    csrc = os.path.join(HERE, "csrc")
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    units = [(os.path.join(csrc, "fx_engine.hip"), os.path.join(bdir, "fx_engine.o"), [])]
    for N in TU_DEGREES:
        for g in TU_GROUPS:
            units.append((os.path.join(csrc, "fx_tu.hip"), os.path.join(bdir, f"fx_tu_n{N}_g{g}.o"),
                          [f"-DFX_TU_N={N}", f"-DFX_TU_GROUP={g}"]))
    flags = [f for f in HIP_FLAGS if f != "-shared"]

    def compile_one(u):
        src, obj, defs = u
        _run([HIPCC, *flags, *defs, "-c", src, "-o", obj + ".tmp"], verbose)
        os.replace(obj + ".tmp", obj)
        return obj

    jobs = jobs or max(1, min(len(units), os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, units))
    out = os.path.join(HERE, "libfluxo_gpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    return out


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)




FORTRAN_SRC = ["fxh_basis.f90", "fxh_mesh.f90", "fluxo_gpu_bind.f90", "fxh_readini.f90", "fxh_dg.f90"]
FLANG = os.environ.get("FLANG", "flang")
F_FLAGS = ["-O2", "-fPIC"]


def build_host(verbose=False):
    """Fortran host: libfluxo_host.so (mesh/metrics + FLUXO-shaped wrappers) and the
    parameter.ini driver fluxo_gpu, both linked against libfluxo_gpu.so."""
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for f in FORTRAN_SRC + ["fluxo_gpu_main.f90"]:
        o = os.path.join(bdir, f.replace(".f90", ".o"))
        _run([FLANG, "-c", *F_FLAGS, "-J", bdir, "-I", bdir, os.path.join(fdir, f), "-o", o], verbose)
        objs.append(o)
    rpath = "-Wl,-rpath,$ORIGIN"
    so = os.path.join(HERE, "libfluxo_host.so")
    _run([FLANG, "-shared", "-o", so + ".tmp", *objs[:-1], "-L" + HERE, "-lfluxo_gpu", "-lfluxo_io", rpath], verbose)
    os.replace(so + ".tmp", so)
    exe = os.path.join(HERE, "fluxo_gpu")
    _run([FLANG, "-o", exe + ".tmp", *objs, "-L" + HERE, "-lfluxo_gpu", "-lfluxo_io", rpath], verbose)
    os.replace(exe + ".tmp", exe)


HDF5_PREFIX = os.environ.get("HDF5_PREFIX", "/opt/conda")


def build_io(verbose=False):
    """libfluxo_io.so: HOPR mesh and state files over the HDF5 C library (host code)."""
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError(f"HDF5 headers not found under {inc} (set HDF5_PREFIX)")
    out = os.path.join(HERE, "libfluxo_io.so")
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I" + inc, os.path.join(HERE, "io", "fx_io.cpp"),
          "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib, "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose=False):
    build_engine(verbose)
    build_io(verbose)
    build_host(verbose)
    build_oracle(verbose)
