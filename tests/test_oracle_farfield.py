"""Pins of the oracle EVALQ / SCAT (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:770-1002)."""
import math

import numpy as np
import pytest

from oracle import ddscat_oracle as O
from ddscat_hip import targets as T


def test_reference_final_section_identities():
    """Restates the reference's numpy self-checks (scripts/test_gpu_final_sections_math.py:3-33)
    against the oracle's own functions."""
    rng = np.random.default_rng(12345)
    N = 257
    r = rng.normal(size=(N, 3))
    p = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    k = rng.normal(size=3)
    k *= 2.3 / np.linalg.norm(k)
    n = k / np.linalg.norm(k)
    phase = np.exp(-1j * (r @ k))
    per_dipole = np.sum((p - n[None, :] * np.sum(p * n[None, :], axis=1)[:, None]) * phase[:, None], axis=0)
    s = np.sum(p * phase[:, None], axis=0)
    assert np.allclose(per_dipole, s - n * np.dot(n, s), rtol=2e-13, atol=2e-13)
    em = rng.normal(size=3)
    em -= n * np.dot(n, em)
    em /= np.linalg.norm(em)
    assert np.allclose(np.sum((p @ em) * phase), np.dot(em, s), rtol=2e-13, atol=2e-13)
    # EVALQ block algebra: component-major compact dot equals the per-dipole sum
    ad = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    ao = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    kk = np.array([1.7, 0.0, 0.0]) ** (1 / 3)  # |k|^3 = 1.7
    e = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    flat = lambda a: a.T.reshape(-1)
    cabs, cext, cpha = O.evalq(flat(e), flat(p), flat(ad), flat(ao), kk, 1.0)
    rabs = np.float64(np.float32(np.linalg.norm(kk) ** 3 / 1.5))
    q = np.empty_like(p)
    q[:, 0] = (ad[:, 0] + 1j * rabs) * p[:, 0] + ao[:, 1] * p[:, 2] + ao[:, 2] * p[:, 1]
    q[:, 1] = (ad[:, 1] + 1j * rabs) * p[:, 1] + ao[:, 2] * p[:, 0] + ao[:, 0] * p[:, 2]
    q[:, 2] = (ad[:, 2] + 1j * rabs) * p[:, 2] + ao[:, 0] * p[:, 1] + ao[:, 1] * p[:, 0]
    z2 = sum(np.vdot(p[j], q[j]) for j in range(N))
    akk = np.linalg.norm(kk)
    assert abs(cabs - (-4 * math.pi * akk * z2.imag)) < 1e-10 * abs(cabs)
    zep = sum(np.vdot(e[j], p[j]) for j in range(N))
    assert abs(cext - 4 * math.pi * akk * zep.imag) < 1e-10 * abs(cext)
    assert abs(cpha - 2 * math.pi * akk * zep.real) < 1e-10 * abs(cpha)


@pytest.mark.parametrize("etasca", [0.5, 1.0, 3.0])
def test_quadrature_weights_cover_the_sphere(etasca):
    dirs, w = O.scat_integration_dirs([0.3, 0.0, 0.0], [0, 1, 0], etasca, 5000)
    assert abs(w[:, 0].sum() - 4 * math.pi) < 1e-12
    k = np.linalg.norm(dirs, axis=1)
    assert np.allclose(k, 0.3, rtol=1e-6)
    assert np.allclose(dirs[0] / 0.3, [1, 0, 0], atol=1e-6) and np.allclose(dirs[-1] / 0.3, [-1, 0, 0], atol=1e-6)
    # <cos> and <cos^2> moments: 0 and 4 pi / 3 up to a quadrature error that
    # grows with the angular step ETASCA
    assert abs(w[:, 1].sum()) < 0.03 * etasca and abs(w[:, 4].sum() - 4 * math.pi / 3) < 0.06 * etasca


def test_single_dipole_radiated_power():
    """One dipole P at a lattice site: Csca = (8 pi / 3) k^4 |P|^2 (closed form)."""
    k = 0.4
    P = np.array([0.3 + 0.1j, 1.0 - 0.2j, 0.05j])
    occ = np.array([0])
    res = O.scat(P, occ, 1, 1, [k, 0, 0], (1, 1, 1), (-1, -1, -1), [0, 1, 0], 0.1, 1.0)
    exact = 8 * math.pi / 3 * k ** 4 * np.vdot(P, P).real
    assert abs(res["csca"] - exact) < 2e-3 * exact
    # backscatter: |E_perp|^2 at theta = pi is |P_y|^2 + |P_z|^2
    assert abs(res["cbksca"] - k ** 4 * (abs(P[1]) ** 2 + abs(P[2]) ** 2)) < 1e-10


def mie_q(x, m):
    """Mie Q_ext, Q_sca of a homogeneous sphere (series of Bohren & Huffman,
    logarithmic-derivative downward recurrence)."""
    mx = m * x
    nstop = int(x + 4 * x ** (1 / 3) + 2)
    nmx = int(max(nstop, abs(mx))) + 16
    D = np.zeros(nmx + 1, complex)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / mx - 1 / (D[n] + n / mx)
    psi0, psi1 = math.cos(x), math.sin(x)
    chi0, chi1 = -math.sin(x), math.cos(x)
    xi1 = complex(psi1, -chi1)
    qext = qsca = 0.0
    for n in range(1, nstop + 1):
        psi = (2 * n - 1) / x * psi1 - psi0
        chi = (2 * n - 1) / x * chi1 - chi0
        xi = complex(psi, -chi)
        an = ((D[n] / m + n / x) * psi - psi1) / ((D[n] / m + n / x) * xi - xi1)
        bn = ((m * D[n] + n / x) * psi - psi1) / ((m * D[n] + n / x) * xi - xi1)
        qext += (2 * n + 1) * (an + bn).real
        qsca += (2 * n + 1) * (abs(an) ** 2 + abs(bn) ** 2)
        psi0, psi1, chi0, chi1, xi1 = psi1, psi, chi1, chi, xi
    return 2 / x ** 2 * qext, 2 / x ** 2 * qsca


def test_mie_series_known_values():
    """Textbook values: m=1.5, x=1 gives Q_ext = Q_sca = 0.2151 (non-absorbing);
    large-x extinction paradox tends to Q_ext -> 2."""
    qe, qs = mie_q(1.0, 1.5 + 0j)
    assert abs(qe - 0.21505) < 2e-4 and abs(qe - qs) < 1e-10
    assert abs(mie_q(300.0, 1.33 + 0.01j)[0] - 2.0) < 0.05


def test_dda_sphere_against_mie_and_optical_theorem():
    """Oracle pipeline (LDR alpha, incident field, dense direct solve, EVALQ,
    SCAT) on a D=12 sphere: Q_ext within DDA accuracy of Mie; Cext = Cabs + Csca."""
    m, xs = 1.33 + 0.05j, 1.0
    tg = T.sphere(12)
    kd = T.kd_for_size_parameter(tg, xs)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = np.array([0, 1, 0], complex)
    adia, aoff = T.operators(tg, m, kd, khat, e0)
    x0 = tg.centroid_x0()
    b = T.incident_field(tg, khat * kd, e0, x0).astype(complex)
    A = O.dense_operator(tg.nx, tg.ny, tg.nz, kd, adia.astype(complex), aoff.astype(complex), tg.iocc)
    occ3 = np.flatnonzero(np.tile(tg.iocc != 0, 3))
    p = np.linalg.solve(A[np.ix_(occ3, occ3)], b[occ3])
    occ = tg.occ
    comp = lambda v: np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)
    cabs, cext, cpha = O.evalq(comp(b), p, comp(adia), comp(aoff), khat * kd, 1.0)
    sc = O.scat(p, occ, tg.nx, tg.ny, khat * kd, (1, 1, 1), x0, e0, 0.5, 1.0)
    pia2 = math.pi * tg.a_eff() ** 2
    qe_mie, qs_mie = mie_q(xs, m)
    assert abs(cext / pia2 - qe_mie) < 0.05 * qe_mie
    assert abs(sc["csca"] / pia2 - qs_mie) < 0.06 * qs_mie
    assert abs(cext - cabs - sc["csca"]) < 0.01 * cext
