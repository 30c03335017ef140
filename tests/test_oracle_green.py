"""Pins of the oracle's Green tensor (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:610-652)
by independent mathematics, not by re-typing its formula."""
import math

import numpy as np
import pytest
import scipy.special as sp

from oracle import ddscat_oracle as O


@pytest.mark.parametrize("x", [1e-6, 0.05, 0.7, 1.99, 2.0, 2.01, 3.3, 7.5, 40.0, 300.0, -0.3, -5.0])
def test_cisi_matches_scipy(x):
    ci, si = O.cisi(x)
    s_ref, c_ref = sp.sici(abs(x))
    assert abs(ci - c_ref) < 1e-13 * max(1.0, abs(c_ref))
    assert abs(si - math.copysign(s_ref, x)) < 1e-13


def dipole_field(r, p, k):
    """Radiating point dipole (textbook form, e.g. Jackson 9.18):
    E = e^{ikr} [k^2 (n x p) x n / r + (3 n (n.p) - p)(1/r^3 - i k / r^2)]."""
    rr = np.linalg.norm(r)
    n = r / rr
    return np.exp(1j * k * rr) * (k * k * np.cross(np.cross(n, p), n) / rr
                                  + (3 * n * np.dot(n, p) - p) * (1 / rr ** 3 - 1j * k / rr ** 2))


@pytest.mark.parametrize("seed", range(5))
def test_green_is_the_dipole_field_propagator(seed):
    rng = np.random.default_rng(seed)
    r = rng.normal(size=3) * 3
    p = rng.normal(size=3) + 1j * rng.normal(size=3)
    k = rng.uniform(0.05, 1.5)
    g = O.green_elements(*r, k)
    G = np.empty((3, 3), complex)
    for c, (i, j) in enumerate(O.COMP_IJ):
        G[i, j] = G[j, i] = g[c]
    assert np.allclose(G @ p, dipole_field(r, p, k), rtol=1e-12, atol=0)


def test_green_zero_at_origin_and_even_in_displacement():
    assert np.all(O.green_elements(0.0, 0.0, 0.0, 0.3) == 0)
    a = O.green_elements(1.3, -2.0, 0.7, 0.4)
    b = O.green_elements(-1.3, 2.0, -0.7, 0.4)
    assert np.array_equal(a, b)


def test_fcd_reduces_to_point_dipole_without_filter_terms():
    rng = np.random.default_rng(3)
    x, y, z = rng.normal(size=(3, 7)) * 4
    pd = O.green_elements(x, y, z, 0.37, 0)
    fc = O.fcd_without_filter(x, y, z, 0.37)
    assert np.allclose(fc, pd, rtol=1e-12, atol=0)


def test_fcd_approaches_point_dipole_far_away():
    """Filter terms (Ci/Si of (pi +- k) r) decay: relative difference ~ 1/r."""
    d1 = np.abs(O.green_elements(3.0, 1.0, 2.0, 0.2, 1) - O.green_elements(3.0, 1.0, 2.0, 0.2, 0)).max()
    d2 = np.abs(O.green_elements(30.0, 10.0, 20.0, 0.2, 1) - O.green_elements(30.0, 10.0, 20.0, 0.2, 0)).max()
    ref2 = np.abs(O.green_elements(30.0, 10.0, 20.0, 0.2, 0)).max()
    assert d2 / ref2 < 0.2 and d2 < d1


def brute_spectral(nx, ny, nz, akd, dx, kx, ky, kz):
    """Direct sum over all lags (-N < l < N) with exp(+i pi k l / N): no FFT,
    no parity argument, no octant trimming."""
    out = np.zeros(6, complex)
    for lz in range(-nz + 1, nz):
        for ly in range(-ny + 1, ny):
            for lx in range(-nx + 1, nx):
                g = O.green_elements(lx * dx[0], ly * dx[1], lz * dx[2], akd)
                ph = np.exp(1j * math.pi * (kx * lx / nx + ky * ly / ny + kz * lz / nz))
                out += g * ph
    return out


def test_octant_matches_brute_force_spectral_sum():
    nx, ny, nz, akd, dx = 3, 4, 2, 0.41, (1.0, 1.1, 1 / 1.1)
    go = O.green_octant(nx, ny, nz, akd, dx)
    for (kx, ky, kz) in [(0, 0, 0), (1, 2, 1), (3, 4, 2), (2, 1, 0), (3, 0, 1)]:
        assert np.allclose(go[:, kz, ky, kx], brute_spectral(nx, ny, nz, akd, dx, kx, ky, kz), rtol=1e-12, atol=1e-12)


def test_full_reconstruction_signs_match_direct_transform():
    """The octant + parity-sign rebuild (cu:109-115) equals the untrimmed transform."""
    nx, ny, nz, akd = 4, 3, 5, 0.3
    g = O.green_lag_grid(nx, ny, nz, akd)
    full = np.fft.ifftn(g, axes=(1, 2, 3)) * (8 * nx * ny * nz)
    rebuilt = O.green_full_from_octant(O.green_octant(nx, ny, nz, akd), nx, ny, nz)
    assert np.allclose(rebuilt, full, rtol=1e-12, atol=1e-12)


def test_separable_parity_transform_equals_fft_octant():
    """The device build's algorithm (lag octant + per-axis cos/sin sums) gives
    the reference's FFT-and-trim octant."""
    import bench
    for dims in [(5, 4, 6), (8, 8, 8)]:
        a = bench._green_octant_fast(*dims, 0.27)
        b = O.green_octant(*dims, 0.27)
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
