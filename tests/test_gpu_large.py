"""Full-size MATVEC checks at the large-target shape (sphere of diameter 384
dipoles, 29.6 M dipoles, padded 768^3 x 3 grid) through size-independent
properties, and an oracle comparison at 96^3 (padded 192, LDS-staged
kernels).

A = alpha^-1 - G is complex symmetric, so y^T (A x) = x^T (A y); 'C' mode is
conj(A conj(x)); the MATVEC is linear; the FFT3D and SLICES backends compute
the same operator.  Bars: symmetry and backend agreement within 1e-5
relative (FP32 MATVEC, FP64 sums), linearity within 1e-5 relative L2, 'C'
identity bit-exact."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from ddscat_hip import targets as T

pytestmark = pytest.mark.gpu


def cvec(rng, n, scale=1.0):
    return (scale * (rng.normal(size=n) + 1j * rng.normal(size=n))).astype(np.complex64)


def apply_compact(lib, cwhat, x, ad, ao):
    import torch
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    xd, add, aod = dev(x), dev(ad), dev(ao)
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    lib.apply_device(cwhat, xd.data_ptr(), yd.data_ptr(), add.data_ptr(), aod.data_ptr(), layout=1)
    lib.sync()
    y = yd.cpu().numpy()
    del xd, add, aod, yd
    torch.cuda.empty_cache()
    return y


def test_generic_kernels_96_match_oracle(lib):
    nx = ny = nz = 96
    rng = np.random.default_rng(96)
    iocc = (rng.random(nx * ny * nz) < 0.5).astype(np.int16)
    occ = np.flatnonzero(iocc)
    akd = 0.2
    lib.prepare_green(nx, ny, nz, iocc, (0.0, akd, 0.0))
    n0 = 3 * occ.size
    x, ad, ao = cvec(rng, n0), cvec(rng, n0), cvec(rng, n0, 0.2)
    y = apply_compact(lib, "N", x, ad, ao)
    full = lambda v: np.zeros(3 * nx * ny * nz, complex)
    xf, adf, aof = full(0), full(0), full(0)
    for f, c in ((xf, x), (adf, ad), (aof, ao)):
        f.reshape(3, -1)[:, occ] = c.reshape(3, -1)
    yref = O.matvec_fft("N", xf, adf, aof, O.green_octant(nx, ny, nz, float(np.float32(akd))), nx, ny, nz, iocc)
    yref = yref.reshape(3, -1)[:, occ].reshape(-1)
    err = np.linalg.norm(y - yref) / np.linalg.norm(yref)
    print(f"96^3 generic MATVEC rel L2 = {err:.2e}")
    assert err < 1e-5


@pytest.fixture(scope="module")
def big():
    tg = T.sphere(384)
    kd = 0.25
    return tg, kd


@pytest.fixture(scope="module")
def products():
    """A x of each backend on the same inputs, shared by the tests of this module."""
    return {}


def sphere384_inputs(tg, kd):
    rng = np.random.default_rng(384)
    n0 = 3 * tg.nat0
    x, y = cvec(rng, n0), cvec(rng, n0)
    m = 1.7 + 0.05j
    alpha_inv = (1.0 / T.ldr_alpha(np.array([m * m]), kd, (1, 0, 0), (0, 1, 0)))[0]
    ad = np.full(n0, alpha_inv, np.complex64)
    ao = cvec(rng, n0, 0.05)
    return x, y, ad, ao


def sphere384_product(lib, tg, kd, backend, x, ad, ao):
    from ddscat_hip import BACKEND_FFT3D, BACKEND_SLICES
    lib.set_backend(BACKEND_SLICES if backend == "slices" else BACKEND_FFT3D)
    try:
        lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (kd, 0.0, 0.0))
        return apply_compact(lib, "N", x, ad, ao)
    finally:
        lib.set_backend(BACKEND_FFT3D)
        lib.release()


@pytest.mark.parametrize("backend", ["fft3d", "slices"])
def test_sphere384_matvec_properties(lib, big, products, backend):
    from ddscat_hip import BACKEND_FFT3D, BACKEND_SLICES
    tg, kd = big
    lib.set_backend(BACKEND_SLICES if backend == "slices" else BACKEND_FFT3D)
    try:
        lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (kd, 0.0, 0.0))
        x, y, ad, ao = sphere384_inputs(tg, kd)
        ax = apply_compact(lib, "N", x, ad, ao)
        ay = apply_compact(lib, "N", y, ad, ao)
        lhs = np.dot(y.astype(complex), ax.astype(complex))
        rhs = np.dot(x.astype(complex), ay.astype(complex))
        sym = abs(lhs - rhs) / abs(lhs)
        # 'C' mode identity (conjugation is exact in every kernel)
        cx = apply_compact(lib, "C", x, ad, ao)
        cref = np.conj(apply_compact(lib, "N", np.conj(x), ad, ao))
        # linearity
        a, bcoef = np.complex64(0.3 - 1.2j), np.complex64(2.0 + 0.5j)
        lin = apply_compact(lib, "N", (a * x + bcoef * y).astype(np.complex64), ad, ao)
        lref = a * ax + bcoef * ay
        lerr = np.linalg.norm(lin - lref) / np.linalg.norm(lref)
        print(f"{backend} 384^3 sphere ({tg.nat0} dipoles): symmetry {sym:.2e}, linearity {lerr:.2e}")
        assert sym < 1e-5
        assert np.array_equal(cx, cref)
        assert lerr < 1e-5
        products[backend] = ax
    finally:
        lib.set_backend(BACKEND_FFT3D)
        lib.release()


def test_sphere384_backends_agree(lib, big, products):
    """FFT3D and SLICES give the same A x over all 89 M entries (products of
    the property tests when they ran in this module, computed here otherwise)."""
    tg, kd = big
    x, _, ad, ao = sphere384_inputs(tg, kd)
    for backend in ("fft3d", "slices"):
        if backend not in products:
            products[backend] = sphere384_product(lib, tg, kd, backend, x, ad, ao)
    ya, yb = products["fft3d"], products["slices"]
    assert ya.shape == yb.shape == (3 * tg.nat0,)
    err = np.linalg.norm(ya - yb) / np.linalg.norm(ya)
    worst = max(np.linalg.norm(ya[c::3] - yb[c::3]) / np.linalg.norm(ya[c::3]) for c in range(3))
    print(f"FFT3D vs SLICES at 384^3 ({ya.size} entries): relative L2 {err:.2e}, worst stride-3 slice {worst:.2e}")
    assert err < 1e-5 and worst < 1e-5
