"""Compute calls through the reference's own loader.

The reference's Fortran bridge reaches the plugin through
ddscat_cuda_unified/cuda/ddscat_cuda_loader.c (dlopen of DDSCAT_CUDA_DLL,
ld:41-102), built here unchanged from its source by oracle/Makefile into
oracle/_ref/.  A D=16 sphere goes through the whole path the DDSCAT driver
takes: ddscat_cuda_loader_prepare_green_f32 -> apply_f32 ('N' and 'C') ->
solve_pbcgs2_f32 -> evalq_f32 -> scat_f32 -> release, all host buffers in
the full [3][NAT] layout, each result checked against the FP64 oracle:
MATVEC within 1e-5 relative L2, Q_ext / Q_abs of the solve within 1e-4 of
the oracle's 1e-9 solution, C_sca and the selected amplitudes of the device
solution within 1e-6 of the oracle's sums over the same polarization.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np
import pytest

from oracle import ddscat_oracle as O
from oracle import solvers as S
from ddscat_hip import targets as T, LIB_FFT3D

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
REF_LOADER = ROOT / "oracle" / "_ref" / "libddscat_cuda_loader.so"


def p(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.fixture(scope="module")
def loader(lib):
    if not REF_LOADER.exists():
        pytest.skip("oracle/_ref/libddscat_cuda_loader.so not built (oracle/Makefile needs the reference source)")
    os.environ["DDSCAT_CUDA_DLL"] = str(LIB_FFT3D)
    L = C.CDLL(str(REF_LOADER))
    f, i, vp = C.c_float, C.c_int, C.c_void_p
    L.ddscat_cuda_loader_prepare_green_f32.argtypes = [i] * 7 + [f] * 9 + [vp]
    L.ddscat_cuda_loader_apply_f32.argtypes = [i, vp, vp, vp, vp]
    L.ddscat_cuda_loader_solve_pbcgs2_f32.argtypes = [vp] * 4 + [i, i, f] + [vp] * 3
    L.ddscat_cuda_loader_evalq_f32.argtypes = [i, vp, vp, vp, vp, f, f, f, f, vp, vp, vp]
    L.ddscat_cuda_loader_scat_f32.argtypes = [i, i, vp, vp, vp, vp, vp, f, f, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.ddscat_cuda_loader_last_error.restype = C.c_char_p
    yield L
    L.ddscat_cuda_loader_release()


def test_sphere_through_reference_loader(loader):
    L = loader
    tg = T.sphere(16)
    kd = T.kd_for_size_parameter(tg, 2.0)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = np.array([0, 1, 0], complex)
    m = 1.33 + 0.01j
    adia, aoff = T.operators(tg, m, kd, khat, e0)
    x0 = tg.centroid_x0()
    b = T.incident_field(tg, khat * kd, e0, x0)
    ak = (khat * kd).astype(np.float32)
    akd32 = float(np.linalg.norm(ak.astype(np.float64)))
    iocc = np.ascontiguousarray(tg.iocc, np.int16)
    rc = L.ddscat_cuda_loader_prepare_green_f32(tg.nx, tg.ny, tg.nz, 0, tg.nat, tg.nat0, 0, 0.0, 0.0, 0.0,
                                                *map(float, ak), 1.0, 1.0, 1.0, p(iocc))
    assert rc == 0, L.ddscat_cuda_loader_last_error()
    # MATVEC 'N' and 'C' against the oracle
    go = O.green_octant(tg.nx, tg.ny, tg.nz, akd32)
    rng = np.random.default_rng(5)
    x = ((rng.normal(size=3 * tg.nat) + 1j * rng.normal(size=3 * tg.nat)) * np.tile(tg.iocc != 0, 3)).astype(
        np.complex64)
    for cw in ("N", "C"):
        y = np.zeros_like(x)
        assert L.ddscat_cuda_loader_apply_f32(ord(cw), p(x), p(y), p(adia), p(aoff)) == 0
        ref = O.matvec_fft(cw, x, adia, aoff, go, tg.nx, tg.ny, tg.nz, tg.iocc)
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        print(f"loader MATVEC '{cw}': relative L2 {err:.2e}")
        assert err <= 1e-5
    # PBCGS2 solve, EVALQ on the resident solution
    xs = np.zeros_like(b)
    it, tole, nc = C.c_int(0), C.c_float(0.0), C.c_int(0)
    rc = L.ddscat_cuda_loader_solve_pbcgs2_f32(p(b), p(xs), p(adia), p(aoff), 3 * tg.nat, 300, 1e-5, C.byref(it),
                                               C.byref(tole), C.byref(nc))
    assert rc == 0 and tole.value <= 1e-5, (rc, tole.value)
    cabs, cext, cpha = C.c_double(), C.c_double(), C.c_double()
    rc = L.ddscat_cuda_loader_evalq_f32(1, None, None, None, None, *map(float, ak), 1.0, C.byref(cabs), C.byref(cext),
                                        C.byref(cpha))
    assert rc == 0
    occ = tg.occ
    comp = lambda v: np.asarray(v).reshape(3, -1)[:, occ].reshape(-1).astype(np.complex128)
    op = O.FftMatvec64(go, tg.nx, tg.ny, tg.nz, tg.iocc, workers=8)
    bc, adc, aoc = comp(b), comp(adia), comp(aoff)
    xo, _, _, _ = S.qmrccg(lambda cw, v: op(cw, v, adc, aoc), bc, maxit=500, tol=1e-9)
    qa, qe, _ = O.evalq(bc, xo, adc, aoc, ak, 1.0)
    print(f"loader PBCGS2: {it.value} iterations, residual {tole.value:.2e}; Cext {cext.value:.8f} (oracle {qe:.8f}),"
          f" Cabs {cabs.value:.8f} (oracle {qa:.8f})")
    assert abs(cext.value - qe) <= 1e-4 * abs(qe) and abs(cabs.value - qa) <= 1e-4 * abs(qa)
    # SCAT over the quadrature plus selected directions, against the oracle's sums of the same solution
    th = np.radians([0.0, 30.0, 90.0, 150.0, 180.0])
    aks = (np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], 1) * kd).astype(np.float32)
    em1 = np.stack([-np.sin(th), np.cos(th), np.zeros_like(th)], 1).astype(np.float32)
    em2 = np.tile(np.float32([0, 0, 1]), (len(th), 1))
    dx = np.ones(3, np.float32)
    x0f = x0.astype(np.float32)
    e01 = e0.astype(np.complex64)
    cbk, csca, cg2 = C.c_double(), C.c_double(), C.c_double()
    cg = np.zeros(3)
    navg = C.c_int(0)
    f1 = np.zeros(len(th), np.complex64)
    f2 = np.zeros(len(th), np.complex64)
    rc = L.ddscat_cuda_loader_scat_f32(0, len(th), p(ak), p(aks), p(dx), p(em1), p(em2), 1.0, 0.5, p(e01), p(x0f),
                                       C.byref(cbk), C.byref(csca), p(cg), C.byref(cg2), C.byref(navg), p(f1), p(f2))
    assert rc == 0, L.ddscat_cuda_loader_last_error()
    ref = O.scat(comp(xs), occ, tg.nx, tg.ny, ak.astype(np.float64), (1.0, 1.0, 1.0), x0f.astype(np.float64),
                 e01.astype(np.complex128), 0.5, 1.0, aks=aks, em1=em1, em2=em2)
    print(f"loader SCAT: Csca {csca.value:.8f} (oracle {ref['csca']:.8f}), {navg.value} directions")
    assert navg.value == ref["navg"]
    assert abs(csca.value - ref["csca"]) <= 1e-6 * ref["csca"]
    assert np.allclose(f1, ref["f1"], rtol=1e-5, atol=1e-6 * np.abs(ref["f1"]).max())
    assert np.allclose(f2, ref["f2"], rtol=1e-5, atol=1e-6 * np.abs(ref["f1"]).max())
    assert L.ddscat_cuda_loader_release() == 0
