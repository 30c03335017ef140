"""Amplitude scattering matrix and Mueller matrix of the far-field output path.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product library.

The reference driver collects, per scattering direction, the amplitudes of
the two incident polarizations from ddscat_cuda_scat_f32 (F1 = f_1j,
F2 = f_2j for incident polarization j; cu:841-845, DDSCAT_cuda.f90:3250-3260)
and turns them into the 2x2 amplitude matrix and the 4x4 Mueller matrix in
GETMUELLER (DDSCAT_cuda.f90:3294-3300).  GETMUELLER and the scattering-basis
routine SCAVEC are not part of the reference repository, so this module
restates the published definitions instead (Bohren & Huffman, "Absorption
and Scattering of Light by Small Particles", 1983, ch. 3):

  * far field of the dipole array: E_s . e_i = e^{ikr}/(kr) f_ij E0 with
    f_ij = k^3 e_i . S_j / |E0|, S_j = sum_l P_l^(j) exp(-i k_s . r_l)
    (the amplitudes the reference's SCAT returns);
  * Bohren & Huffman: (E_par_s, E_perp_s) = e^{ikr}/(-ikr) [[S2, S3], [S4, S1]]
    (E_par_i, E_perp_i), with e_perp = e_perp_i = e_perp_s normal to the
    scattering plane, e_par = khat x e_perp for the incident and the
    scattered wave;
  * hence [[S2, S3], [S4, S1]] = -i F M, F = [[f11, f12], [f21, f22]] in the
    bases (e_par_s, e_perp) <- (e01, e02), M_jk = conj(e0j) . (e_par_i, e_perp)_k;
  * Mueller matrix: Bohren & Huffman eq. 3.16.

Pins (tests/test_oracle_mueller.py): the Mueller formula against Stokes
vectors propagated through the amplitude matrix (B&H eq. 2.84 definitions);
a single point dipole against the small-sphere limit of the Mie series
(fixes the -i factor and the par/perp assignment); Mie S1, S2 against the
optical theorem and their forward equality; the DDA sphere's Mueller matrix
against Mie at DDA accuracy.  The relation to DDSCAT's own SCAVEC basis
vectors is parity unpinned (SCAVEC is not vendored).
"""
from __future__ import annotations

import math
import numpy as np


def directions(khat, e01, theta, phi):
    """Scattering directions and bases for incidence ``khat``: phi is measured
    from a1 = Re(e01) (normalised, the same frame as the reference's SCAT
    quadrature, cu:849-875), a2 = khat x a1.
    Returns (ks [n,3] unit vectors, e_par_s [n,3], e_perp [n,3])."""
    khat = np.asarray(khat, float) / np.linalg.norm(khat)
    a1 = np.real(np.asarray(e01, complex))
    a1 = a1 - khat * np.dot(a1, khat)
    a1 = a1 / np.linalg.norm(a1)
    a2 = np.cross(khat, a1)
    th = np.asarray(theta, float).reshape(-1, 1)
    ph = np.asarray(phi, float).reshape(-1, 1)
    ks = np.cos(th) * khat + np.sin(th) * (np.cos(ph) * a1 + np.sin(ph) * a2)
    e_perp = np.sin(ph) * a1 - np.cos(ph) * a2
    e_par = np.cross(ks, e_perp)
    return ks, e_par, e_perp


def amplitudes(Sj, ak, e_par_s, e_perp, e02=1.0):
    """f_1j, f_2j of one incident polarization from its far-field sums
    Sj [n,3] (cu:841-845): f = k^3/|E0| (e . S)."""
    k = float(np.linalg.norm(ak))
    term = k ** 3 / math.sqrt(e02)
    return term * np.einsum("nk,nk->n", e_par_s, Sj), term * np.einsum("nk,nk->n", e_perp, Sj)


def amplitude_matrix(f11, f21, f12, f22, e01, e02v, khat, e_perp):
    """Bohren & Huffman S1..S4 from the f_ij of incident polarizations e01,
    e02v (orthonormal complex 3-vectors) for scattering planes with normal
    e_perp [n,3]; returns (S1, S2, S3, S4)."""
    khat = np.asarray(khat, float) / np.linalg.norm(khat)
    e_perp = np.asarray(e_perp, float)
    e_par_i = np.cross(khat, e_perp)
    c01, c02 = np.conj(np.asarray(e01, complex)), np.conj(np.asarray(e02v, complex))
    m11, m12 = e_par_i @ c01, e_perp @ c01
    m21, m22 = e_par_i @ c02, e_perp @ c02
    S2 = -1j * (f11 * m11 + f12 * m21)
    S3 = -1j * (f11 * m12 + f12 * m22)
    S4 = -1j * (f21 * m11 + f22 * m21)
    S1 = -1j * (f21 * m12 + f22 * m22)
    return S1, S2, S3, S4


def mueller(S1, S2, S3, S4):
    """4x4 Mueller matrix [..., 4, 4] of the amplitude matrix (B&H eq. 3.16)."""
    S1, S2, S3, S4 = (np.asarray(v, complex) for v in (S1, S2, S3, S4))
    a1, a2, a3, a4 = (np.abs(v) ** 2 for v in (S1, S2, S3, S4))
    M = np.empty(S1.shape + (4, 4))
    M[..., 0, 0] = 0.5 * (a2 + a3 + a4 + a1)
    M[..., 0, 1] = 0.5 * (a2 - a3 + a4 - a1)
    M[..., 0, 2] = (S2 * np.conj(S3) + S1 * np.conj(S4)).real
    M[..., 0, 3] = (S2 * np.conj(S3) - S1 * np.conj(S4)).imag
    M[..., 1, 0] = 0.5 * (a2 - a4 + a3 - a1)
    M[..., 1, 1] = 0.5 * (a2 - a3 - a4 + a1)
    M[..., 1, 2] = (S2 * np.conj(S3) - S1 * np.conj(S4)).real
    M[..., 1, 3] = (S2 * np.conj(S3) + S1 * np.conj(S4)).imag
    M[..., 2, 0] = (S2 * np.conj(S4) + S1 * np.conj(S3)).real
    M[..., 2, 1] = (S2 * np.conj(S4) - S1 * np.conj(S3)).real
    M[..., 2, 2] = (S1 * np.conj(S2) + S3 * np.conj(S4)).real
    M[..., 2, 3] = (S2 * np.conj(S1) + S4 * np.conj(S3)).imag
    M[..., 3, 0] = (np.conj(S2) * S4 + np.conj(S3) * S1).imag
    M[..., 3, 1] = (np.conj(S2) * S4 - np.conj(S3) * S1).imag
    M[..., 3, 2] = (S1 * np.conj(S2) - S3 * np.conj(S4)).imag
    M[..., 3, 3] = (S1 * np.conj(S2) - S3 * np.conj(S4)).real
    return M


def stokes(Epar, Eperp):
    """Stokes vector (I, Q, U, V) of a field with components (E_par, E_perp)
    (Bohren & Huffman eq. 2.84)."""
    I = abs(Epar) ** 2 + abs(Eperp) ** 2
    Q = abs(Epar) ** 2 - abs(Eperp) ** 2
    U = (Epar * np.conj(Eperp) + Eperp * np.conj(Epar)).real
    V = (1j * (Epar * np.conj(Eperp) - Eperp * np.conj(Epar))).real
    return np.array([I, Q, U, V])


def mie_s12(x, m, theta):
    """Mie amplitudes S1(theta), S2(theta) of a homogeneous sphere (size
    parameter x, relative index m): Bohren & Huffman's series with the
    angular functions pi_n, tau_n and a_n, b_n from the logarithmic
    derivative (downward recurrence)."""
    theta = np.atleast_1d(np.asarray(theta, float))
    mu = np.cos(theta)
    mx = m * x
    nstop = int(x + 4 * x ** (1 / 3) + 2)
    nmx = int(max(nstop, abs(mx))) + 16
    D = np.zeros(nmx + 1, complex)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / mx - 1 / (D[n] + n / mx)
    psi0, psi1 = math.cos(x), math.sin(x)
    chi0, chi1 = -math.sin(x), math.cos(x)
    xi1 = complex(psi1, -chi1)
    S1 = np.zeros_like(mu, complex)
    S2 = np.zeros_like(mu, complex)
    pi0, pi1 = np.zeros_like(mu), np.ones_like(mu)
    for n in range(1, nstop + 1):
        psi = (2 * n - 1) / x * psi1 - psi0
        chi = (2 * n - 1) / x * chi1 - chi0
        xi = complex(psi, -chi)
        an = ((D[n] / m + n / x) * psi - psi1) / ((D[n] / m + n / x) * xi - xi1)
        bn = ((m * D[n] + n / x) * psi - psi1) / ((m * D[n] + n / x) * xi - xi1)
        pin = pi1
        taun = n * mu * pin - (n + 1) * pi0
        fn = (2 * n + 1) / (n * (n + 1))
        S1 += fn * (an * pin + bn * taun)
        S2 += fn * (an * taun + bn * pin)
        pi0, pi1 = pin, ((2 * n + 1) * mu * pin - (n + 1) * pi0) / n
        psi0, psi1, chi0, chi1, xi1 = psi1, psi, chi1, chi, xi
    return S1, S2
