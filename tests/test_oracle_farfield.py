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


def _rings(dirs, ak):
    """theta of every direction and the distinct ring angles (FP32 directions)."""
    kh = np.asarray(ak, float) / np.linalg.norm(ak)
    par = dirs @ kh
    perp = np.linalg.norm(dirs - par[:, None] * kh[None, :], axis=1)
    th = np.arctan2(perp, par)  # accurate next to the poles, unlike arccos
    rings = np.unique(np.round(th, 6))
    return th, rings


@pytest.mark.parametrize("x_target, etasca", [(2.0, 0.5), (10.0, 0.25), (10.0, 0.1), (48.0, 0.5)])
def test_quadrature_ring_spacing_rule(x_target, etasca):
    """The ring count and theta map (cu:849-875) put the last ring spacing,
    next to theta = pi, at (pi/2) eta / (3 + x) and the first one, next to
    the forward direction, at theta0 / (1 + theta0) (1 + theta0 / (pi +
    theta0)^2) of that (theta0 = 2 pi / (1 + x), finer where the diffraction
    peak is).  Both follow from
    the map's derivative at its ends, theta'(s=0) = theta0 / (1 + theta0)
    and theta'(s_max) = 1 / (1 + theta0 / (pi + theta0)^2), with the ring
    count 1 + round(2 (3 + x) / eta * (1 + 1/(pi + theta0)) /
    (1 + theta0/(pi + theta0)^2)).  A wrong end point, a mis-scaled map or
    a ring count off by a few fails; the exact count within one is the
    reference's definitional rounding (cu:849-875), which no rule of the
    geometry fixes."""
    nat0 = 20000
    kd = x_target / (3 * nat0 / (4 * math.pi)) ** (1 / 3)
    ak = [kd, 0.0, 0.0]
    dirs, w = O.scat_integration_dirs(ak, [0, 1, 0], etasca, nat0)
    th, rings = _rings(dirs, ak)
    nth = len(rings)
    x = kd * (3 * nat0 / (4 * math.pi)) ** (1 / 3)
    theta0 = 2 * math.pi / (1 + x)
    back = (math.pi / 2) * etasca / (3 + x)
    assert rings[0] == 0.0 and abs(rings[-1] - math.pi) < 1e-5
    # each end spacing is the map's slope there times the s step ds, plus
    # the curvature over one step (relative size below ds at both ends) and
    # the rounding of the ring count (relative 1 / n_theta)
    ds = (math.pi + math.pi / (theta0 + math.pi)) / (nth - 1)
    slack = 2.0 / nth + ds
    assert abs((rings[-1] - rings[-2]) / back - 1) < slack, (rings[-1] - rings[-2], back)
    fwd = back * theta0 / (1 + theta0) * (1 + theta0 / (math.pi + theta0) ** 2)
    assert abs((rings[1] - rings[0]) / fwd - 1) < slack, (rings[1] - rings[0], fwd)
    # each interior ring carries n_phi = round(2 pi sin(theta) / dtheta) >= 3
    # equally weighted azimuths (cells about as wide as the ring spacing
    # dtheta, half the distance between its neighbours): the ring's weights
    # sum to its solid-angle band
    for i in (1, nth // 2, nth - 2):
        sel = np.abs(th - rings[i]) < 2e-6
        nphi = int(sel.sum())
        assert nphi >= 3
        band = 2 * math.pi * (math.cos(0.5 * (rings[i - 1] + rings[i])) - math.cos(0.5 * (rings[i] + rings[i + 1])))
        assert abs(w[sel, 0].sum() - band) < 1e-3 * band
        dth = 0.5 * (rings[i + 1] - rings[i - 1])
        assert abs(nphi - 2 * math.pi * math.sin(rings[i]) / dth) <= 0.5 + 1e-3 or nphi == 3


@pytest.mark.parametrize("etasca", [1.0, 0.5, 0.25])
def test_quadrature_integrates_low_harmonics(etasca):
    """Integrals over the sphere of the degree <= 4 functions the far-field
    sums meet (the dipole pattern 1 - (n.a1)^2, cos^2, cos^4, and odd or
    azimuthal terms that must vanish) converge to their exact values at the
    rate of a midpoint rule in theta, O(dtheta^2): halving ETASCA cuts the
    error about four times."""
    nat0 = 5000
    ak = np.array([0.4, 0.0, 0.0])
    dirs, w = O.scat_integration_dirs(ak, [0, 1, 0], etasca, nat0)
    n = dirs / np.linalg.norm(dirs, axis=1)[:, None]
    ct, a1c, a2c = n[:, 0], n[:, 1], n[:, 2]
    wt = w[:, 0]
    exact = {
        "dipole": (lambda: 1 - a1c ** 2, 8 * math.pi / 3),
        "cos2": (lambda: ct ** 2, 4 * math.pi / 3),
        "cos4": (lambda: ct ** 4, 4 * math.pi / 5),
        "cos": (lambda: ct, 0.0),
        "a1a2": (lambda: a1c * a2c, 0.0),
        "a1": (lambda: a1c, 0.0),
        "a2^2-a1^2": (lambda: a2c ** 2 - a1c ** 2, 0.0),
    }
    errs = {k: abs(float(np.sum(wt * f())) - v) for k, (f, v) in exact.items()}
    # the step ~ (pi/2) eta / (3 + x) sets the bound
    x = 0.4 * (3 * nat0 / (4 * math.pi)) ** (1 / 3)
    h = (math.pi / 2) * etasca / (3 + x)
    for k, e in errs.items():
        assert e < 4 * math.pi * h * h, (k, e, h)
    # the weight columns are the moments of the direction cosines
    assert np.allclose(w[:, 1], wt * ct, atol=1e-7)
    assert np.allclose(w[:, 4], wt * ct ** 2, atol=1e-7)
    assert np.allclose(w[:, 2] ** 2 + w[:, 3] ** 2, (wt ** 2) * (1 - ct ** 2), atol=1e-7)
