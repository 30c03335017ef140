"""Pins of the amplitude-matrix / Mueller restatement (oracle/scattering.py).

GETMUELLER is not part of the reference repository; the restatement follows
Bohren & Huffman and is fixed here by:
  * the Mueller formula against Stokes vectors propagated through the
    amplitude matrix (independent of eq. 3.16's algebra);
  * Mie S1, S2: equal in the forward direction, optical theorem
    Q_ext = 4 Re S(0) / x^2 against a textbook Q_ext;
  * a single point dipole (the DDA far field of one site) against the
    small-sphere limit of the Mie series: fixes the -i factor and which
    amplitude is parallel / perpendicular;
  * the full DDA chain on a D=12 sphere (dense solve of both incident
    polarizations, far-field sums, amplitudes, Mueller) against Mie at DDA
    accuracy, and the block structure a sphere imposes (S3 = S4 = 0)."""
import math

import numpy as np
import pytest

from oracle import ddscat_oracle as O
from oracle import scattering as SC
from ddscat_hip import targets as T


def test_mueller_formula_propagates_stokes_vectors():
    rng = np.random.default_rng(3)
    for _ in range(5):
        S1, S2, S3, S4 = rng.normal(size=4) + 1j * rng.normal(size=4)
        M = SC.mueller(S1, S2, S3, S4)
        for _ in range(6):
            Ep, Es = rng.normal(size=2) + 1j * rng.normal(size=2)
            out = SC.stokes(S2 * Ep + S3 * Es, S4 * Ep + S1 * Es)
            assert np.allclose(M @ SC.stokes(Ep, Es), out, rtol=1e-12, atol=1e-12)


def test_mie_amplitudes_forward_and_optical_theorem():
    S1, S2 = SC.mie_s12(1.0, 1.5 + 0j, [0.0])
    assert abs(S1[0] - S2[0]) < 1e-14
    assert abs(4 * S1[0].real / 1.0 ** 2 - 0.21505) < 2e-4   # Q_ext(m=1.5, x=1), non-absorbing
    # Rayleigh angular law |S2|/|S1| = |cos theta| for a tiny sphere
    th = np.linspace(0.1, 3.0, 7)
    S1, S2 = SC.mie_s12(0.01, 1.5 + 0.1j, th)
    assert np.allclose(np.abs(S2 / S1), np.abs(np.cos(th)), rtol=1e-3)


def test_point_dipole_matches_small_sphere():
    x, m = 0.02, 1.5 + 0.1j
    k = 1.0
    a = x / k
    alpha = a ** 3 * (m * m - 1) / (m * m + 2)
    khat = np.array([1.0, 0.0, 0.0])
    e01 = np.array([0, 1, 0], complex)
    e02 = np.cross(khat, e01)
    th = np.linspace(0.0, math.pi, 9)
    for phi in (0.0, 0.7, 2.0):
        ks, epar, eperp = SC.directions(khat, e01, th, np.full_like(th, phi))
        f = []
        for e0 in (e01, e02):
            Sj = np.tile(alpha * e0, (th.size, 1))       # P = alpha E0, one site at the origin
            f.append(SC.amplitudes(Sj, k * khat, epar, eperp))
        (f11, f21), (f12, f22) = f
        S1, S2, S3, S4 = SC.amplitude_matrix(f11, f21, f12, f22, e01, e02, khat, eperp)
        m1, m2 = SC.mie_s12(x, m, th)
        assert np.allclose(S1, m1, rtol=2e-3, atol=0)
        assert np.allclose(S2, m2, rtol=2e-3, atol=1e-3 * abs(m1[0]))
        assert np.max(np.abs(S3)) < 1e-12 * abs(m1[0]) and np.max(np.abs(S4)) < 1e-12 * abs(m1[0])


def test_amplitude_matrix_rejects_swapped_bases():
    """A dipole seen with the parallel and perpendicular scattered bases swapped
    no longer matches the sphere: the pin above is not vacuous."""
    x, m = 0.02, 1.5 + 0.1j
    alpha = x ** 3 * (m * m - 1) / (m * m + 2)
    khat = np.array([1.0, 0.0, 0.0])
    e01 = np.array([0, 1, 0], complex)
    e02 = np.cross(khat, e01)
    th = np.array([0.5, 1.0, 2.0])
    ks, epar, eperp = SC.directions(khat, e01, th, np.zeros(3))
    f = [SC.amplitudes(np.tile(alpha * e0, (3, 1)), khat, eperp, epar) for e0 in (e01, e02)]
    S = SC.amplitude_matrix(f[0][0], f[0][1], f[1][0], f[1][1], e01, e02, khat, eperp)
    m1, m2 = SC.mie_s12(x, m, th)
    assert np.max(np.abs(SC.mueller(*S) - SC.mueller(m1, m2, 0 * m1, 0 * m1))) > 0.1 * np.max(np.abs(m1)) ** 2


@pytest.fixture(scope="module")
def dda_sphere():
    m, xs = 1.33 + 0.05j, 1.0
    tg = T.sphere(12)
    kd = T.kd_for_size_parameter(tg, xs)
    khat = np.array([1.0, 0.0, 0.0])
    e01 = np.array([0, 1, 0], complex)
    e02 = np.cross(khat, e01).astype(complex)
    x0 = tg.centroid_x0()
    occ = tg.occ
    comp = lambda v: np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)
    occ3 = np.flatnonzero(np.tile(tg.iocc != 0, 3))
    P = []
    for e0 in (e01, e02):
        adia, aoff = T.operators(tg, m, kd, khat, e0)
        b = T.incident_field(tg, khat * kd, e0, x0).astype(complex)
        A = O.dense_operator(tg.nx, tg.ny, tg.nz, kd, adia.astype(complex), aoff.astype(complex), tg.iocc)
        P.append(np.linalg.solve(A[np.ix_(occ3, occ3)], b[occ3]))
    return dict(tg=tg, kd=kd, khat=khat, e01=e01, e02=e02, x0=x0, P=P, m=m, xs=xs, comp=comp)


def test_dda_sphere_mueller_matches_mie(dda_sphere):
    D = dda_sphere
    tg = D["tg"]
    th = np.radians(np.arange(0, 181, 15.0))
    res = {}
    for phi in (0.0, 0.5 * math.pi, 1.1):
        ks, epar, eperp = SC.directions(D["khat"], D["e01"], th, np.full_like(th, phi))
        f = [SC.amplitudes(O.farfield_sums(p, tg.occ, tg.nx, tg.ny, (1, 1, 1), D["x0"], ks * D["kd"]),
                           D["khat"] * D["kd"], epar, eperp) for p in D["P"]]
        S = SC.amplitude_matrix(f[0][0], f[0][1], f[1][0], f[1][1], D["e01"], D["e02"], D["khat"], eperp)
        res[phi] = SC.mueller(*S)
    # Mie, with the DDA sphere's own a_eff
    x = D["kd"] * tg.a_eff()
    m1, m2 = SC.mie_s12(x, D["m"], th)
    Mm = SC.mueller(m1, m2, np.zeros_like(m1), np.zeros_like(m1))
    for phi, M in res.items():
        s11 = M[:, 0, 0]
        assert np.all(np.abs(s11 / Mm[:, 0, 0] - 1) < 0.06), (phi, s11 / Mm[:, 0, 0])
        for (i, j) in ((0, 1), (2, 2), (2, 3), (3, 3)):
            assert np.all(np.abs(M[:, i, j] / s11 - Mm[:, i, j] / Mm[:, 0, 0]) < 0.03), (phi, i, j)
        # a sphere's amplitude matrix is diagonal: the off-block elements vanish
        for (i, j) in ((0, 2), (0, 3), (1, 2), (1, 3), (2, 0), (3, 0)):
            assert np.all(np.abs(M[:, i, j]) < 0.02 * s11), (phi, i, j)
