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


def filtered_tensor_by_quadrature(x, y, z, k, kf=math.pi):
    """The filtered-coupled-dipole interaction from its definition, with no
    Ci/Si closed form (Piller & Martin 1998; Gay-Balmaz & Martin 2002; Yurkin,
    Min & Hoekstra 2010): the spectrum of the scalar Green function e^{ikr}/r,
    4 pi / (q^2 - k^2), cut at q < k_F, gives

        g_F(r) = (2 / pi) PV int_0^kF q / (q^2 - k^2) sin(q r) / r dq + i sin(k r) / r

    (the pole at q = k gives the outgoing imaginary part), the band-limited
    delta function is delta_F(r) = (1 / (2 pi^2 r)) int_0^kF q sin(q r) dq, and

        G_F = (k^2 I + grad grad) g_F + (4 pi / 3) delta_F I.

    The radial derivatives are quadratures of d/dr, d^2/dr^2 of sin(qr)/r, and
    grad grad f(r) = f'' n n + f'/r (I - n n).  Returns the six components
    xx, xy, xz, yy, yz, zz."""
    from scipy.integrate import quad
    r = math.sqrt(x * x + y * y + z * z)

    def pv(h):
        # PV int_0^kF q h(q) / (q^2 - k^2) dq = PV int q h(q) / (q + k) / (q - k)
        return quad(lambda q: q * h(q) / (q + k), 0.0, kf, weight="cauchy", wvar=k, limit=400,
                    epsabs=1e-14, epsrel=1e-12)[0]

    s0 = lambda q: math.sin(q * r) / r
    s1 = lambda q: q * math.cos(q * r) / r - math.sin(q * r) / r ** 2
    s2 = lambda q: -q * q * math.sin(q * r) / r - 2 * q * math.cos(q * r) / r ** 2 + 2 * math.sin(q * r) / r ** 3
    i0 = math.sin(k * r) / r
    i1 = k * math.cos(k * r) / r - math.sin(k * r) / r ** 2
    i2 = -k * k * math.sin(k * r) / r - 2 * k * math.cos(k * r) / r ** 2 + 2 * math.sin(k * r) / r ** 3
    f0 = 2 / math.pi * pv(s0) + 1j * i0
    f1 = 2 / math.pi * pv(s1) + 1j * i1
    f2 = 2 / math.pi * pv(s2) + 1j * i2
    delta_f = quad(lambda q: q * math.sin(q * r), 0.0, kf, limit=200, epsabs=1e-15)[0] / (2 * math.pi ** 2 * r)
    n = np.array([x, y, z]) / r
    G = (k * k * f0 + f1 / r + 4 * math.pi / 3 * delta_f) * np.eye(3) + (f2 - f1 / r) * np.outer(n, n)
    return np.array([G[i, j] for i, j in O.COMP_IJ])


@pytest.mark.parametrize("k", [0.05, 0.3, 1.1])
def test_fcd_matches_band_limited_spectral_definition(k):
    """Every filter term of the FCD tensor (cu:627-631: the Ci/Si combinations
    a0, a1, a2 and the smeared delta h2) against quadratures of the
    band-limited spectrum.  A dropped or sign-flipped term, or the cut at a
    wrong k_F, moves some component by far more than the 1e-8 tolerance."""
    rng = np.random.default_rng(int(k * 100))
    pts = [(1.0, 0.0, 0.0), (1.0, 1.0, 0.0), (1.0, 1.0, 1.0), (2.0, -1.0, 3.0)] + \
        [tuple(v) for v in rng.integers(-9, 10, size=(4, 3)) if np.any(v)]
    for p in pts:
        want = filtered_tensor_by_quadrature(*map(float, p), k)
        got = O.green_elements(*map(float, p), k, 1)
        assert np.abs(got - want).max() <= 1e-8 * np.abs(want).max(), (p, got, want)
        # and the filter matters at these distances: the point-dipole tensor is far off
        pd = O.green_elements(*map(float, p), k, 0)
        if math.sqrt(sum(v * v for v in p)) < 4:
            assert np.abs(pd - want).max() > 1e-3 * np.abs(want).max()


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
    for dims in [(5, 4, 6), (8, 8, 8)]:
        a = O.green_octant_separable(*dims, 0.27)
        b = O.green_octant(*dims, 0.27)
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
