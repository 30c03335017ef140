"""GPU parity of the SLICES (low-memory) MATVEC backend: same operator as
FFT3D (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:361-404), checked
against the FP64 oracle at the 1e-5 relative-L2 MATVEC bar and against the
FFT3D backend on the device."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from ddscat_hip import Library, BACKEND_FFT3D, BACKEND_SLICES, targets as T

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.complex128) - b) / np.linalg.norm(b))


def random_case(rng, nx, ny, nz, fill=0.7):
    nat = nx * ny * nz
    iocc = (rng.random(nat) < fill).astype(np.int16)
    iocc[0] = 1
    mask = np.tile(iocc != 0, 3)
    cx = lambda: (rng.normal(size=3 * nat) + 1j * rng.normal(size=3 * nat)).astype(np.complex64)
    return iocc, cx() * mask, cx(), cx() * 0.3


@pytest.fixture()
def slib(lib):
    lib.release()
    lib.set_backend(BACKEND_SLICES)
    yield lib
    lib.release()
    lib.set_backend(BACKEND_FFT3D)


# register-FFT y passes (2 NY = 32, 48, 128, 256, 256, 512 = 2 x 256,
# 768 = 3 x 256), register z passes with the gather and the finish (2 NZ =
# 32, 128, 256, 512, 768), register x passes at 2 NX = 512 and 768 (32 x 16,
# 32 x 24), and all three axes on register kernels
# (16, 24, 32); the others, and the remaining axes, run the LDS-staged kernels
@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 5, 4), (8, 8, 8), (12, 15, 16), (24, 20, 18), (10, 6, 36),
                                  (5, 16, 6), (3, 24, 5), (4, 64, 3), (3, 128, 4), (7, 128, 3), (4, 256, 3),
                                  (3, 384, 4), (6, 5, 16), (5, 7, 64), (3, 4, 128), (3, 3, 256), (2, 3, 384),
                                  (16, 24, 32), (256, 3, 4), (384, 4, 3)])
@pytest.mark.parametrize("cwhat", ["N", "C"])
def test_slices_matvec_matches_oracle(slib, dims, cwhat):
    nx, ny, nz = dims
    rng = np.random.default_rng(sum(dims) * 3 + (cwhat == "C"))
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    akd = 0.29
    slib.prepare_green(nx, ny, nz, iocc, (akd, 0.0, 0.0))
    assert slib.info()["backend"] == BACKEND_SLICES
    y = slib.apply(cwhat, x, adia, aoff)
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)))
    yref = O.matvec_fft(cwhat, x, adia, aoff, goct, nx, ny, nz, iocc)
    err = rel_l2(y, yref)
    print(f"SLICES {dims} {cwhat}: rel L2 = {err:.3e}")
    assert err < 1e-5
    assert np.all(y.reshape(3, -1)[:, iocc == 0] == 0)


def test_slices_lds_kernels_on_register_lengths(slib):
    """ddscat_hip_set_fft_kernels(1) keeps the LDS-staged SLICES kernels on a
    grid whose axes all have register shapes (the register z path's buffer
    is allocated but unused)."""
    nx, ny, nz = 16, 24, 32
    rng = np.random.default_rng(162432)
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(0.29)))
    try:
        slib.set_fft_kernels(1)
        slib.prepare_green(nx, ny, nz, iocc, (0.29, 0.0, 0.0))
        for cw in ("N", "C"):
            err = rel_l2(slib.apply(cw, x, adia, aoff), O.matvec_fft(cw, x, adia, aoff, goct, nx, ny, nz, iocc))
            print(f"SLICES LDS-staged kernels {cw}: {err:.2e}")
            assert err < 1e-5
    finally:
        slib.set_fft_kernels(0)


def test_slices_compact_layout_and_solver(slib):
    """D=16 sphere: every padded axis is 32 points, so the x, y and z passes
    run the register kernels and the compact layout goes through the site-map
    gather and finish of the z passes."""
    import torch
    tg = T.sphere(16)
    kd = T.kd_for_size_parameter(tg, 1.5)
    khat = np.array([0.0, 0.0, 1.0])
    e0 = np.array([1, 0, 0], complex)
    adia, aoff = T.operators(tg, 1.5 + 0.02j, kd, khat, e0)
    b = T.incident_field(tg, khat * kd, e0, tg.centroid_x0())
    slib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    rng = np.random.default_rng(1)
    x = ((rng.normal(size=3 * tg.nat) + 1j * rng.normal(size=3 * tg.nat)) * np.tile(tg.iocc != 0, 3)).astype(np.complex64)
    comp = lambda v: np.ascontiguousarray(v.reshape(3, -1)[:, tg.occ].reshape(-1))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()
    xc, ac, oc = dev(comp(x)), dev(comp(adia)), dev(comp(aoff))
    yc = torch.empty_like(xc)
    for cw in ("N", "C"):
        yfull = slib.apply(cw, x, adia, aoff)
        torch.cuda.synchronize()
        slib.apply_device(cw, xc.data_ptr(), yc.data_ptr(), ac.data_ptr(), oc.data_ptr(), layout=1)
        slib.sync()
        assert rel_l2(yc.cpu().numpy().view(np.complex64), comp(yfull).astype(np.complex128)) < 1e-6, cw
    xs, it, tole, nc, rc = slib.solve("qmrccg", b, adia, aoff, maxit=300, tol=1e-5)
    slib.release()
    slib.set_backend(BACKEND_FFT3D)
    slib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    xf, it2, tole2, nc2, rc2 = slib.solve("qmrccg", b, adia, aoff, maxit=300, tol=1e-5)
    print(f"SLICES qmrccg it={it} tole={tole:.2e}; FFT3D it={it2} tole={tole2:.2e}")
    assert rc == 0 and tole <= 1e-5
    assert rel_l2(xs, xf.astype(np.complex128)) < 3e-5


def test_slice_library_defaults_to_slices():
    from ddscat_hip import LIB_SLICES
    L = Library(LIB_SLICES)
    nx = ny = nz = 4
    L.prepare_green(nx, ny, nz, np.ones(64, np.int16), (0.2, 0, 0))
    assert L.info()["backend"] == BACKEND_SLICES
    rng = np.random.default_rng(2)
    x = (rng.normal(size=192) + 1j * rng.normal(size=192)).astype(np.complex64)
    ad = (rng.normal(size=192) + 1j).astype(np.complex64)
    ao = np.zeros(192, np.complex64)
    y = L.apply("N", x, ad, ao)
    yref = O.matvec_fft("N", x, ad, ao, O.green_octant(nx, ny, nz, float(np.float32(0.2))), nx, ny, nz)
    assert rel_l2(y, yref) < 1e-5
    L.release()
