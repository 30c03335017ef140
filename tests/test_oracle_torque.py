"""Pins of the oracle's radiative-torque sums (oracle.scattering.torque_sums,
DDSCAT's DOTORQ quantities; the reference's GPU SCAT returns them to the CPU,
ddscat_cuda_unified/GPU_FINAL_SECTIONS.md:75-80).

  * an absorbing sphere in circularly polarised light takes up the angular
    momentum of the photons it absorbs: C_Gamma . khat = +-C_abs (Draine &
    Weingartner 1996), and the incident-beam part alone is +-C_ext;
  * a single dipole, the same in closed form;
  * an asymmetric dipole array in elliptically polarised light from an
    oblique direction: the far-field balance equals the torque summed over
    every dipole in the total field (pair fields and their gradients from
    the closed-form tensor, radiation reaction included), which no part of
    torque_sums computes that way.
"""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from oracle import scattering as SC
from ddscat_hip import targets as T


def dense_solution(tg, m, kd, khat, e0, x0):
    ad, ao = T.operators(tg, m, kd, khat, e0)
    A = O.dense_operator(tg.nx, tg.ny, tg.nz, kd, ad.astype(complex), ao.astype(complex), tg.iocc)
    r = T.site_positions_full(tg, x0)
    Ef = np.asarray(e0, complex)[:, None] * np.exp(1j * (r @ (np.asarray(khat) * kd)))[None, :]
    Ef[:, tg.iocc == 0] = 0
    Ef = Ef.reshape(-1)
    hole = np.diag((np.tile(tg.iocc, 3) == 0).astype(float))
    Pf = np.linalg.solve(A + hole, Ef)
    comp = lambda v: np.asarray(v).reshape(3, -1)[:, tg.occ].reshape(-1)
    cabs, cext, _ = O.evalq(comp(Ef), comp(Pf), comp(ad), comp(ao), np.asarray(khat) * kd, 1.0)
    return comp(Pf), comp(Ef), cabs, cext


@pytest.mark.parametrize("helicity", [1, -1])
def test_absorbing_sphere_takes_up_absorbed_angular_momentum(helicity):
    tg = T.sphere(10)
    kd = T.kd_for_size_parameter(tg, 1.5)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = np.array([0.0, 1.0, helicity * 1j]) / np.sqrt(2)
    x0 = tg.centroid_x0()
    P, E, cabs, cext = dense_solution(tg, 1.5 + 0.2j, kd, khat, e0, x0)
    t = SC.torque_sums(P, E, tg.occ, tg.nx, tg.ny, khat * kd, (1, 1, 1), x0, 0.1, e0)
    assert abs(t["ext"][0] - helicity * cext) < 1e-10 * cext
    assert abs(t["total"][0] - helicity * cabs) < 2e-4 * cabs
    assert np.all(np.abs(t["total"][1:]) < 1e-6 * cabs)
    # linear polarisation: no torque on a sphere
    P, E, cabs, _ = dense_solution(tg, 1.5 + 0.2j, kd, khat, np.array([0, 1.0, 0]), x0)
    t = SC.torque_sums(P, E, tg.occ, tg.nx, tg.ny, khat * kd, (1, 1, 1), x0, 0.1, np.array([0, 1.0, 0]))
    assert np.all(np.abs(t["total"]) < 1e-6 * cabs)


def test_single_dipole_closed_form():
    """One dipole, P = alpha E0: C_Gamma = (C_ext - C_sca) khat for helicity +1,
    with C_ext = 4 pi k Im(alpha) and C_sca = (8 pi / 3) k^4 |alpha|^2."""
    kd = 0.7
    alpha = 0.05 + 0.02j
    e0 = np.array([0.0, 1.0, 1j]) / np.sqrt(2)
    P = alpha * e0
    t = SC.torque_sums(P, e0, np.array([0]), 1, 1, (kd, 0, 0), (1, 1, 1), (-1.0, -1.0, -1.0), 0.05, e0, nat0=1)
    want = 4 * np.pi * kd * alpha.imag - 8 * np.pi / 3 * kd ** 4 * abs(alpha) ** 2
    assert abs(t["total"][0] - want) < 1e-4 * abs(want)


def test_far_field_balance_equals_dipole_mechanics():
    rng = np.random.default_rng(3)
    tg = T.Target(5, 4, 4, (rng.random(80) < 0.6).astype(np.int16))
    kd = 0.6
    khat = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    e1, e2 = T.polarization_pair(khat)
    e0 = (e1 + 0.6j * e2) / np.linalg.norm(e1 + 0.6j * e2)
    x0 = np.array([-3.0, -2.5, -2.5])
    P, E, cabs, _ = dense_solution(tg, 1.5 + 0.3j, kd, khat, e0, x0)
    t = SC.torque_sums(P, E, tg.occ, tg.nx, tg.ny, khat * kd, (1, 1, 1), x0, 0.1, e0)
    mech = SC.torque_dense_mechanics(P, E, tg.occ, tg.nx, tg.ny, tg.nz, khat * kd, x0)
    err = np.linalg.norm(t["total"] - mech) / np.linalg.norm(mech)
    assert err < 1e-3, (t["total"], mech)
