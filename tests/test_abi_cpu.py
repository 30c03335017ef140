"""Host-side checks of the drop-in boundary: both libraries load, export
exactly what include/ddscat_matvec_hip.h declares, and answer the host-only
error paths; the reference's own loader (built from its source by
oracle/Makefile) binds them and keeps its contract (1001 / 1002 / -777).
No compute call reaches the GPU here."""
import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

from ddscat_hip import (Library, LIB_FFT3D, LIB_SLICES, REFERENCE_SYMBOLS, EXTENSION_SYMBOLS,
                        exported_symbols)

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "ddscat_matvec_hip.h"
LOADER = ROOT / "tests" / "loader"


def header_symbols():
    text = HEADER.read_text()
    return set(re.findall(r"DDSCAT_HIP_API\s+[\w\s\*]+?\*?\s*(ddscat_\w+)\s*\(", text))


def test_header_declares_reference_and_extension_sets():
    declared = header_symbols()
    assert set(REFERENCE_SYMBOLS) <= declared
    assert set(EXTENSION_SYMBOLS) <= declared
    assert len(REFERENCE_SYMBOLS) == 15


@pytest.mark.parametrize("path", [LIB_FFT3D, LIB_SLICES])
def test_library_exports_every_declared_symbol(path):
    assert path.exists(), "run __graft_entry__.build() first"
    assert exported_symbols(path) >= header_symbols()
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert exported == header_symbols(), "no undeclared exports"


def test_host_side_error_paths_without_device():
    L = Library()
    assert L.release() == 0
    assert L.lib.ddscat_cuda_apply_f32(ord("N"), None, None, None, None) == 10
    assert "not prepared" in L.last_error()
    # invalid grid: nx=0, nat != nx*ny*nz, nat0 > nat, bad IDIPINT
    assert L.lib.ddscat_cuda_prepare_green_f32(0, 2, 2, 0, 8, 8, 0, 0, 0, 0, .1, 0, 0, 1, 1, 1, None) == 201
    assert L.lib.ddscat_cuda_prepare_f32(2, 2, 2, 0, 9, 8, None, None) == 1
    assert L.lib.ddscat_cuda_prepare_green_f32(2, 2, 2, 1, 8, 8, 0, 0.0, 0, 0, .1, 0, 0, 1, 1, 1,
                                               b"\x01\x00" * 8) == 202
    for name, code in {"gpbicg": 30, "qmrccg": 50, "pbcgst": 70, "pbcgs2": 90, "petrkp": 110,
                       "bicgstab4": 130, "gpbicgstab2": 150, "gpbicgstab4": 180}.items():
        fn = getattr(L.lib, f"ddscat_cuda_solve_{name}_f32")
        assert fn(None, None, None, None, 24, 10, 1e-5, None, None, None) == code
    assert L.evalq((0.1, 0, 0), 1.0, check=False)[3] == 300
    assert L.scat((0.1, 0, 0), (1, 1, 1), (0, 0, 0), (0, 1, 0), 0.5, check=False)["rc"] == 320
    assert L.lib.ddscat_hip_set_backend(7) == 1 and L.lib.ddscat_hip_set_precision(3) == 1
    assert L.lib.ddscat_hip_set_polarizability(2) == 1 and L.lib.ddscat_hip_set_polarizability(0) == 0
    assert L.lib.ddscat_hip_solve_log(5, None, None, 0, None, 0, None) == -1
    assert L.lib.ddscat_hip_select_target(-1) == 1 and L.lib.ddscat_hip_select_target(65) == 1
    assert L.lib.ddscat_hip_select_target(0) == 0
    assert L.lib.ddscat_hip_solve_targets(1, 0, None, None, None, None, None, None, 1, 10, 1e-5, None, None, None,
                                          1) == 1
    assert L.info()["nat"] == 0


REF_LOADER = ROOT / "oracle" / "_ref" / "libddscat_cuda_loader.so"

# One loader per process (it keeps the first library it loads): each probe
# runs in a child python that binds the reference's own loader, compiled from
# ddscat_cuda_unified/cuda/ddscat_cuda_loader.c by oracle/Makefile, and makes
# only calls the plugin answers on the host.
PROBE = r"""
import ctypes as C, sys
L = C.CDLL(sys.argv[1])
L.ddscat_cuda_loader_last_error.restype = C.c_char_p
f, i = C.c_float, C.c_int
L.ddscat_cuda_loader_prepare_green_f32.argtypes = [i] * 7 + [f] * 9 + [C.c_void_p]
print("prepare_green_invalid", L.ddscat_cuda_loader_prepare_green_f32(0, 2, 2, 0, 8, 8, 0, 0, 0, 0, .1, 0, 0, 1, 1, 1, None),
      flush=True)
if len(sys.argv) > 2:
    sys.exit(0)
print("apply_unprepared", L.ddscat_cuda_loader_apply_f32(78, None, None, None, None))
L.ddscat_cuda_loader_solve_pbcgs2_f32.argtypes = [C.c_void_p] * 4 + [i, i, f] + [C.c_void_p] * 3
print("pbcgs2_unprepared", L.ddscat_cuda_loader_solve_pbcgs2_f32(None, None, None, None, 24, 10, 1e-5, None, None, None))
print("last_error", L.ddscat_cuda_loader_last_error().decode())
print("release", L.ddscat_cuda_loader_release())
"""


def run_loader(dll, first_call_only=False):
    if not REF_LOADER.exists():
        pytest.skip("oracle/_ref/libddscat_cuda_loader.so not built (needs the reference source; oracle/Makefile)")
    env = dict(os.environ, DDSCAT_CUDA_DLL=str(dll))
    cmd = [sys.executable, "-c", PROBE, str(REF_LOADER)] + (["first"] if first_call_only else [])
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=120).stdout
    return dict(l.split(" ", 1) for l in out.strip().splitlines() if " " in l)


def test_reference_loader_binds_both_libraries():
    for lib in (LIB_FFT3D, LIB_SLICES):
        r = run_loader(lib)
        assert r["prepare_green_invalid"] == "201"
        assert r["apply_unprepared"] == "10"
        assert r["pbcgs2_unprepared"] == "90"
        assert "invalid argument" in r["last_error"]
        assert r["release"] == "0"


def test_reference_loader_failure_codes():
    """The reference loader's own contract: 1001 when dlopen fails, 1002 when
    a required symbol is missing, -777 when only prepare_green is absent (the
    Fortran bridge then builds the Green tensor on the CPU).  After a
    missing-symbol failure the reference loader keeps the library handle and
    a second call goes through a NULL function pointer, so each probe here
    makes one call (the Fortran driver stops at the first nonzero code)."""
    assert run_loader("/nonexistent/libddscat_matvec_cuda.so", True)["prepare_green_invalid"] == "1001"
    subprocess.run(["make", "-C", str(LOADER)], check=True, capture_output=True)
    assert run_loader(LOADER / "_bin" / "libstub_missing.so", True)["prepare_green_invalid"] == "1002"
    assert run_loader(LOADER / "_bin" / "libstub_nogreen.so", True)["prepare_green_invalid"] == "-777"
