"""The C-ABI library loads and exports every symbol include/fluxo_gpu.h declares (no GPU calls)."""
import ctypes
import os
import re

from fluxo_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "fluxo_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fx_[a-z_0-9]+)\s*\(", src)))


def test_header_lists_the_python_binding_exports():
    assert set(declared_symbols()) == set(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_ctx_bytes_is_host_only_and_scales():
    small = native.make_config(3, 8, 24)
    big = native.make_config(3, 64, 192)
    a, b = native.ctx_bytes(small), native.ctx_bytes(big)
    assert 0 < a < b and abs(b / a - 8.0) < 0.2


def io_declared_symbols():
    src = open(os.path.join(ROOT, "include", "fluxo_io.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fxio_[a-z_0-9]+)\s*\(", src)))


def test_io_library_exports_every_declared_symbol():
    from fluxo_amd import io
    lib = ctypes.CDLL(io.IO_LIB)
    names = io_declared_symbols()
    assert len(names) == 7
    for name in names:
        assert hasattr(lib, name), name


def test_libraries_load_alone_with_all_symbols_resolved():
    """Each shared library loads in a fresh process with every symbol bound at load time
    (RTLD_NOW), so a missing link dependency shows here and not on the first call."""
    import subprocess
    import sys
    for so in ("libfluxo_gpu.so", "libfluxo_io.so", "libfluxo_host.so"):
        path = os.path.join(ROOT, "fluxo_amd", so)
        code = f"import ctypes, os; ctypes.CDLL({path!r}, mode=os.RTLD_NOW)"
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
        assert p.returncode == 0, (so, p.stderr[-500:])


def test_noalloc_rejects_sizes_beyond_its_scratch_on_the_host():
    """fx_rhs_noalloc checks the configuration and the scratch it needs before any HIP call:
    a mesh too large for the static scratch array is refused with FX_EINVAL."""
    L = native.lib()
    small = native.make_config(3, 8, 24, parabolic=True)
    big = native.make_config(3, 40 ** 3, 3 * 40 ** 3, parabolic=True)
    limit = 64 * 1024 * 1024
    assert 0 < L.fx_rhs_noalloc_bytes(ctypes.byref(small)) < limit < L.fx_rhs_noalloc_bytes(ctypes.byref(big))
    dummy = (ctypes.c_double * 8)()
    mesh = native.fx_mesh()
    for field, _ in native._MESH_KEYS:
        setattr(mesh, field, ctypes.addressof(dummy))
    rc = L.fx_rhs_noalloc(ctypes.byref(big), ctypes.byref(mesh), ctypes.addressof(dummy), ctypes.addressof(dummy),
                          None)
    assert rc == -1 and b"maximum" in L.fx_last_error()
    mhd = native.make_config(3, 8, 24)
    mhd.nVar, mhd.volumeFlux = 9, 10
    rc = L.fx_rhs_noalloc(ctypes.byref(mhd), ctypes.byref(mesh), ctypes.addressof(dummy), ctypes.addressof(dummy),
                          None)
    assert rc == -1 and b"nVar = 5" in L.fx_last_error()
