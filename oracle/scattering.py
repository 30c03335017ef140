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


# ----------------------------------------------------------------------------
# Radiative torque (DDSCAT's DOTORQ sums; the reference's GPU SCAT hands them
# back to the CPU, ddscat_cuda_unified/GPU_FINAL_SECTIONS.md:75-80, dispatch
# fortran/getfml_cuda.f90:1135-1150)
# ----------------------------------------------------------------------------
def torque_sums(p_c, e_c, occ, nx, ny, ak, dx, x0, etasca, e01, nat0=None):
    """Torque on the dipole array, as torque cross sections C = Gamma k / u_rad
    (u_rad = |E0|^2 / 8 pi, Gaussian units; Q = C / (pi a_eff^2)), in two parts:

      ext: the incident plane wave's torque on the dipoles,
           Gamma = sum_j [ r_j x F_j + (1/2) Re(P_j x E_j^*) ],
           F_j = -(1/2) Im(P_j^* . E_j) k      (E_j the incident field at r_j);
      rad: the recoil of the radiated field, from the radiative part of the
           interaction, G_I(r) = (k^3 / 4 pi) int dOmega (I - n n) e^{i k n.r}
           (which includes the radiation reaction (2/3) k^3 at r = 0):
           Gamma = (k^3 / 4 pi) int dOmega [ (1/2) Re(-i S x S_perp^*)
                                            - (k/2) Re(conj(R) S_perp) x n ],
           S(n) = sum_j P_j e^{-i k n.r_j},  R_ab(n) = sum_j r_ja P_jb e^{-i k n.r_j},
           S_perp = S - n (n.S), integrated with SCAT's ETASCA quadrature.

    The total ext + rad is the torque; for an absorbing sphere in circularly
    polarised light its component along k is +-C_abs.  Returns dict(ext, rad,
    total [3] each, navg).  FP64, compact [3][NAT0] P and E."""
    from .ddscat_oracle import scat_integration_dirs, site_positions
    ak = np.asarray(ak, np.float64)
    k = float(np.linalg.norm(ak))
    e02 = float(np.sum(np.abs(np.asarray(e01, complex)) ** 2))
    n0 = np.asarray(p_c).size // 3
    P = np.asarray(p_c, np.complex128).reshape(3, n0).T
    E = np.asarray(e_c, np.complex128).reshape(3, n0).T
    r = site_positions(occ, nx, ny, dx, x0)
    u = e02 / (8.0 * np.pi)
    # incident-beam torque
    imPE = np.imag(np.sum(np.conj(P) * E, axis=1))                    # Im(P^* . E) per site
    orb = -0.5 * np.cross(np.sum(imPE[:, None] * r, axis=0), ak)     # sum_j r_j x F_j
    spin = 0.5 * np.real(np.sum(np.cross(P, np.conj(E)), axis=0))
    g_ext = orb + spin
    # radiated recoil
    dirs, w = scat_integration_dirs(ak, e01, etasca, n0 if nat0 is None else nat0)
    g_rad = np.zeros(3)
    for i0 in range(0, len(dirs), 256):
        d = dirs[i0:i0 + 256]
        ph = np.exp(-1j * (d @ r.T))                                   # [nd, nat0]
        S = ph @ P                                                     # [nd, 3]
        R = np.einsum("dj,ja,jb->dab", ph, r, P)                       # [nd, 3, 3]
        n = d / k
        Sp = S - n * np.sum(n * S, axis=1, keepdims=True)
        t_spin = 0.5 * np.real(-1j * np.cross(S, np.conj(Sp)))
        t_orb = -0.5 * k * np.cross(np.real(np.einsum("dab,db->da", np.conj(R), Sp)), n)
        g_rad += (k ** 3 / (4 * np.pi)) * np.sum(w[i0:i0 + 256, 0:1] * (t_spin + t_orb), axis=0)
    c = k / u
    return dict(ext=g_ext * c, rad=g_rad * c, total=(g_ext + g_rad) * c, navg=len(dirs))


def torque_dense_mechanics(p_c, e_c, occ, nx, ny, nz, ak, x0, e02=1.0, h=1e-5):
    """Torque cross section from the mechanics of every dipole in the total
    field (test pin of torque_sums, small targets): Gamma = sum_j [ r_j x F_j
    + (1/2) Re(P_j x E_j^*) ], with E_j = E_inc + sum_{l != j} G(r_j - r_l) P_l
    + (2/3) i k^3 P_j (radiation reaction) and F_j = (1/2) Re sum_i P_ji^*
    grad E_i(r_j), the pair-field gradients by central differences of the
    closed-form tensor (step h).  Reactive pair forces cancel in the sum;
    what remains equals torque_sums' ext + rad."""
    from .ddscat_oracle import green_elements, COMP_IJ
    ak = np.asarray(ak, np.float64)
    k = float(np.linalg.norm(ak))
    n0 = np.asarray(p_c).size // 3
    P = np.asarray(p_c, np.complex128).reshape(3, n0).T
    E = np.asarray(e_c, np.complex128).reshape(3, n0).T
    occ = np.asarray(occ)
    ix, iy, iz = occ % nx, (occ // nx) % ny, occ // (nx * ny)
    r = np.stack([ix + 1 + x0[0], iy + 1 + x0[1], iz + 1 + x0[2]], 1).astype(np.float64)

    def tensor(dv):
        g = green_elements(dv[..., 0], dv[..., 1], dv[..., 2], k)
        T = np.zeros(dv.shape[:-1] + (3, 3), np.complex128)
        for c, (a, b) in enumerate(COMP_IJ):
            T[..., a, b] = g[c]
            T[..., b, a] = g[c]
        return T

    dvec = r[:, None, :] - r[None, :, :]                               # r_j - r_l
    field = np.einsum("jlab,lb->ja", tensor(dvec), P)                  # zero for l == j
    Ej = E + field + (2.0 / 3.0) * 1j * k ** 3 * P
    grad = np.zeros((n0, 3, 3), np.complex128)                         # [j, m, i] d_m E_i
    for m in range(3):
        sh = np.zeros(3); sh[m] = h
        off = ~np.eye(n0, dtype=bool)
        gp = np.einsum("jlab,lb->ja", tensor(dvec + sh) * off[..., None, None], P)
        gm = np.einsum("jlab,lb->ja", tensor(dvec - sh) * off[..., None, None], P)
        grad[:, m, :] = (gp - gm) / (2 * h) + 1j * ak[m] * E
    F = 0.5 * np.real(np.einsum("ji,jmi->jm", np.conj(P), grad))
    gam = np.sum(np.cross(r, F), axis=0) + 0.5 * np.real(np.sum(np.cross(P, np.conj(Ej)), axis=0))
    return gam * k / (e02 / (8.0 * np.pi))
