"""The MATVEC at the full sizes of the 256^3 block (16.8 M dipoles, padded
512^3 x 3) and the D=384 sphere (29.6 M dipoles, padded 768^3 x 3) against
FP64 direct sums, with no FFT on the reference side.

x is zero except at S sparse random sites, so (A x)_j at any site j is the
local term plus -sum_s G(r_j - r_s) x_s over those S sites only
(ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:104-118 with the tensor of
cu:622-626): the oracle evaluates it from the closed-form tensor
(oracle.green_elements) at M sampled sites, half of them sources (where the
diagonal and off-diagonal local terms act), half anywhere in the target.  The
sources span the whole lattice, so every pass of the padded FFT pipeline and
lags in every octant up to the full extent enter each sampled value.  Bar:
relative L2 over the sampled entries <= 1e-5 (measured 4e-8 to 5e-8)."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from ddscat_hip import BACKEND_FFT3D, BACKEND_SLICES, targets as T


def sparse_case(tg, n_src, n_out, seed):
    rng = np.random.default_rng(seed)
    n0 = tg.nat0
    src = np.sort(rng.choice(n0, size=n_src, replace=False))
    out = np.unique(np.concatenate([src[rng.choice(n_src, n_out // 2, replace=False)],
                                    rng.choice(n0, size=n_out // 2, replace=False)]))
    cx = lambda n, s=1.0: (s * (rng.normal(size=n) + 1j * rng.normal(size=n))).astype(np.complex64)
    x = np.zeros((3, n0), np.complex64)
    x[:, src] = cx(3 * n_src).reshape(3, n_src)
    ad = cx(3 * n0).reshape(3, n0) + np.complex64(3.0)
    ao = cx(3 * n0, 0.1).reshape(3, n0)
    return src, out, x, ad, ao


def direct_sum(tg, akd, src, out, x, ad, ao, chunk=256):
    """FP64 (A x) at compact sites `out` for x nonzero only at compact sites `src`."""
    occ = tg.occ
    nx, ny = tg.nx, tg.ny
    coords = lambda q: (occ[q] % nx, (occ[q] // nx) % ny, occ[q] // (nx * ny))
    sx, sy, sz = (c.astype(np.float64) for c in coords(src))
    xs = x[:, src].astype(np.complex128)
    y = np.zeros((3, out.size), np.complex128)
    for a in range(0, out.size, chunk):
        o = out[a:a + chunk]
        ox, oy, oz = (c.astype(np.float64) for c in coords(o))
        g = O.green_elements(ox[:, None] - sx[None, :], oy[:, None] - sy[None, :], oz[:, None] - sz[None, :], akd)
        gxx, gxy, gxz, gyy, gyz, gzz = g
        y[0, a:a + chunk] = -(gxx @ xs[0] + gxy @ xs[1] + gxz @ xs[2])
        y[1, a:a + chunk] = -(gxy @ xs[0] + gyy @ xs[1] + gyz @ xs[2])
        y[2, a:a + chunk] = -(gxz @ xs[0] + gyz @ xs[1] + gzz @ xs[2])
    xo = x[:, out].astype(np.complex128)
    ado, (a23, a31, a12) = ad[:, out].astype(np.complex128), ao[:, out].astype(np.complex128)
    y += ado * xo
    y[0] += a31 * xo[2] + a12 * xo[1]
    y[1] += a12 * xo[0] + a23 * xo[2]
    y[2] += a23 * xo[1] + a31 * xo[0]
    return y


def device_product(lib, tg, kd, x, ad, ao):
    import torch
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.reshape(-1))).cuda()
    xd, add, aod = dev(x), dev(ad), dev(ao)
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    lib.apply_device("N", xd.data_ptr(), yd.data_ptr(), add.data_ptr(), aod.data_ptr(), layout=1)
    lib.sync()
    y = yd.cpu().numpy().reshape(3, -1)
    del xd, add, aod, yd
    torch.cuda.empty_cache()
    return y


def check(lib, tg, kd, backend, seed, label):
    src, out, x, ad, ao = sparse_case(tg, n_src=2048, n_out=512, seed=seed)
    lib.release()
    lib.set_backend(BACKEND_SLICES if backend == "slices" else BACKEND_FFT3D)
    try:
        lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (kd, 0.0, 0.0))
        y = device_product(lib, tg, kd, x, ad, ao)
    finally:
        lib.release()
        lib.set_backend(BACKEND_FFT3D)
    ref = direct_sum(tg, float(np.float32(kd)), src, out, x, ad, ao)
    got = y[:, out].astype(np.complex128)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"{label} {backend}: {out.size} sampled sites x 3 vs FP64 direct sums over {src.size} sources: "
          f"rel L2 {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("d", [10, 12])
def test_direct_sum_matches_oracle_fft_matvec(d):
    """CPU: the direct sums equal the oracle's complex128 FFT MATVEC on small
    targets (block 10^3 and sphere D=12), so the reference side of the GPU
    tests below has the MATVEC's conventions (signs, off-diagonal layout,
    compact order)."""
    tg = T.block(10) if d == 10 else T.sphere(12)
    src, out, x, ad, ao = sparse_case(tg, n_src=64, n_out=40, seed=3)
    akd = 0.31
    ref = direct_sum(tg, akd, src, out, x, ad, ao)
    full = lambda v: O.scatter_compact(v.reshape(-1).astype(np.complex128), tg.occ, tg.nat)
    goct = O.green_octant(tg.nx, tg.ny, tg.nz, akd)
    y = np.asarray(O.matvec_fft("N", full(x), full(ad), full(ao), goct, tg.nx, tg.ny, tg.nz, tg.iocc))
    got = y.reshape(3, -1)[:, tg.occ][:, out]
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["fft3d", "slices"])
def test_block256_matvec_matches_direct_sums(lib, backend):
    tg = T.block(256)
    check(lib, tg, 0.1259, backend, 256, "block 256^3")


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["fft3d", "slices"])
def test_sphere384_matvec_matches_direct_sums(lib, backend):
    tg = T.sphere(384)
    check(lib, tg, 0.25, backend, 384, "sphere D=384")
