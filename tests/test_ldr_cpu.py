"""Pins of the harness's LDR polarizabilities (targets.ldr_alpha, the torch
map adjoint.ldr_alpha_inv, and the same formulas in the device setup
ddscat_hip_setup_rhs, checked against these on the GPU).  In DDSCAT they come
from ALPHADIAG (ddscat_cuda_unified/fortran/getfml_cuda.f90:600, not
vendored) with CALPHA = LATTDR or GKDLDR (reapar_cuda.f90:497-513).

Checks that a plausible mistake fails:
  * kd -> 0 gives Clausius-Mossotti exactly (a wrong alpha_CM fails);
  * for real eps the dipole neither absorbs nor gains energy: Im(1/alpha) =
    -(2/3)(kd)^3, i.e. Cext = 4 pi k Im(alpha) equals Csca = (8 pi/3) k^4
    |alpha|^2 (a wrong sign or factor in the radiative term fails);
  * the (kd)^2 term for incidences with S != 0 (DG93 eq. for S =
    sum (a_j e_j)^2), and its difference from an S = 0 incidence;
  * GKDLDR's diagonal elements projected on a real polarization give the
    LATTDR value (sum_j e_j^2 a_j^2 = S), while the element along the
    propagation direction differs.
"""
import math

import numpy as np
import pytest

from ddscat_hip import targets as T

EPS = [2.25 + 0.0j, 1.7 ** 2 + 0.0j, (1.5 + 0.1j) ** 2, 3.9 + 0.05j]


def inv_cm(eps):
    return (4.0 * math.pi / 3.0) * (eps + 2.0) / (eps - 1.0)


@pytest.mark.parametrize("prescription", ["LATTDR", "GKDLDR"])
def test_clausius_mossotti_limit(prescription):
    rng = np.random.default_rng(1)
    for eps in EPS:
        k = rng.normal(size=3)
        a = T.ldr_alpha(np.array([eps]), 1e-7, k, T.polarization_pair(k)[0], prescription)
        want = (3.0 / (4.0 * math.pi)) * (eps - 1.0) / (eps + 2.0)
        assert np.allclose(a, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("prescription", ["LATTDR", "GKDLDR"])
@pytest.mark.parametrize("kd", [0.05, 0.25, 0.6])
def test_lossless_dipole_scatters_what_it_extinguishes(prescription, kd):
    """Optical theorem of a point dipole with real eps: Cext = Csca."""
    rng = np.random.default_rng(int(kd * 100))
    for eps in (2.25 + 0j, 1.7 ** 2 + 0j, 6.0 + 0j):
        k = rng.normal(size=3)
        for e in T.polarization_pair(k):
            a = np.atleast_1d(T.ldr_alpha(np.array([eps]), kd, k, e, prescription)).reshape(-1)
            cext = 4 * math.pi * kd * a.imag
            csca = 8 * math.pi / 3 * kd ** 4 * np.abs(a) ** 2
            assert np.allclose(cext, csca, rtol=1e-10, atol=0)
            assert np.allclose((1.0 / a).imag, -(2.0 / 3.0) * kd ** 3, rtol=1e-10)


def test_s_dependence_of_lattdr():
    """S = 1/2 for incidence along (1,1,0)/sqrt2 polarised along (1,-1,0)/sqrt2,
    S = 0 along a lattice axis; the (kd)^2 term carries b3 eps S."""
    kd, eps = 0.2, 2.25 + 0.3j
    k1, e1 = np.array([1.0, 1.0, 0.0]), np.array([1.0, -1.0, 0.0]) / math.sqrt(2)
    k0, e0 = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
    assert np.allclose(T.ldr_s(k1, e1), 0.5) and np.allclose(T.ldr_s(k0, e0), 0.0)
    rad = -(2.0 / 3.0) * 1j * kd ** 3
    for k, e, S in ((k1, e1, 0.5), (k0, e0, 0.0), (np.array([1.0, 1.0, 1.0]), np.array([1.0, 1.0, -2.0]), 1 / 3)):
        ainv = 1.0 / T.ldr_alpha(np.array([eps]), kd, k, e)[0]
        term = (ainv - inv_cm(eps) - rad) / kd ** 2
        assert abs(term - (T.LDR_B1 + T.LDR_B2 * eps + T.LDR_B3 * S * eps)) < 1e-10
    d = 1.0 / T.ldr_alpha(np.array([eps]), kd, k1, e1)[0] - 1.0 / T.ldr_alpha(np.array([eps]), kd, k0, e0)[0]
    assert abs(d - 0.5 * T.LDR_B3 * eps * kd ** 2) < 1e-12


def test_gkdldr_projects_to_lattdr():
    rng = np.random.default_rng(5)
    for _ in range(6):
        k = rng.normal(size=3)
        for e in T.polarization_pair(k):
            e = e.real
            for eps in EPS:
                g = 1.0 / T.ldr_alpha(np.array([eps]), 0.3, k, e, "GKDLDR")[:, 0]
                d = 1.0 / T.ldr_alpha(np.array([eps]), 0.3, k, e, "LATTDR")[0]
                assert abs(np.sum(e ** 2 * g) - d) < 1e-12 * abs(d)
                khat = k / np.linalg.norm(k)
                # along the propagation direction: S_j = khat_j^2 sums to 1, not S
                along = np.sum(khat ** 2 * g)
                assert abs(along - d) > 1e-4 * abs(d)


def test_torch_ldr_map_matches_host():
    torch = pytest.importorskip("torch")
    from ddscat_hip.adjoint import ldr_alpha_inv
    rng = np.random.default_rng(9)
    eps = 1.0 + 3.0 * rng.random(50) + 0.05j
    k = np.array([0.2, -0.5, 0.8])
    e = T.polarization_pair(k)[1]
    got = ldr_alpha_inv(torch.tensor(eps), 0.17, k, e).numpy()
    want = 1.0 / T.ldr_alpha(eps, 0.17, k, e)
    assert np.allclose(got, want, rtol=1e-12, atol=0)
