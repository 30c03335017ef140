"""Device radiative torque (ddscat_hip_torque; DDSCAT's DOTORQ sums, which the
reference's GPU SCAT hands back to the CPU: GPU_FINAL_SECTIONS.md:75-80).

  * against the FP64 oracle (oracle.scattering.torque_sums, pinned by dipole
    mechanics in test_oracle_torque.py) on the device's own solution, for an
    asymmetric target in elliptically polarised light from an oblique
    direction: incident-beam and radiated parts within 1e-6 relative;
  * the absorbing-sphere pin through the device chain: in circularly
    polarised light the torque along k equals C_abs from EVALQ (2e-3, the
    ETASCA quadrature's accuracy), the transverse components vanish, and the
    opposite helicity reverses it.
"""
import numpy as np
import pytest

from oracle import scattering as SC
from ddscat_hip import targets as T

pytestmark = pytest.mark.gpu


def solve(lib, tg, m, kd, khat, e0, x0, tol=1e-6):
    ad, ao = T.operators(tg, m, kd, khat, e0)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, np.asarray(khat) * kd)
    b = T.incident_field(tg, np.asarray(khat) * kd, e0, x0)
    x, it, tole, nc, rc = lib.solve("pbcgs2", b, ad, ao, maxit=500, tol=tol)
    comp = lambda v: np.asarray(v).reshape(3, -1)[:, tg.occ].reshape(-1)
    return comp(x), comp(b)


def test_torque_matches_oracle(lib):
    rng = np.random.default_rng(12)
    tg = T.Target(9, 8, 7, (rng.random(9 * 8 * 7) < 0.7).astype(np.int16))
    kd = 0.4
    khat = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    e1, e2 = T.polarization_pair(khat)
    e0 = ((e1 + 0.6j * e2) / np.linalg.norm(e1 + 0.6j * e2)).astype(np.complex64)
    x0 = np.array([-5.0, -4.5, -4.0])
    P, E = solve(lib, tg, 1.6 + 0.2j, kd, khat, e0, x0)
    ak = (khat * kd).astype(np.float32)
    got = lib.torque(ak, e0, 0.5, x0=x0)
    want = SC.torque_sums(P, E, tg.occ, tg.nx, tg.ny, ak.astype(np.float64), (1, 1, 1), x0.astype(np.float32),
                          0.5, e0.astype(np.complex128))
    print(f"device ext {got['ext']} rad {got['rad']}; oracle ext {want['ext']} rad {want['rad']}")
    assert got["navg"] == want["navg"]
    for part in ("ext", "rad"):
        assert np.linalg.norm(got[part] - want[part]) <= 1e-6 * np.linalg.norm(want[part]), part


@pytest.mark.parametrize("helicity", [1, -1])
def test_absorbing_sphere_torque_is_absorbed_angular_momentum(lib, helicity):
    tg = T.sphere(16)
    kd = T.kd_for_size_parameter(tg, 2.0)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = (np.array([0.0, 1.0, helicity * 1j]) / np.sqrt(2)).astype(np.complex64)
    x0 = tg.centroid_x0()
    solve(lib, tg, 1.5 + 0.2j, kd, khat, e0, x0)
    ak = (khat * kd).astype(np.float32)
    cabs, cext, _, _ = lib.evalq(ak)
    t = lib.torque(ak, e0, 0.2, x0=x0)
    print(f"helicity {helicity}: C_Gamma {t['total']} (ext {t['ext'][0]:.6f}, rad {t['rad'][0]:.6f}), "
          f"C_abs {cabs:.6f}, C_ext {cext:.6f}, {t['navg']} directions")
    assert abs(t["ext"][0] - helicity * cext) <= 1e-5 * cext
    assert abs(t["total"][0] - helicity * cabs) <= 2e-3 * cabs
    assert np.all(np.abs(t["total"][1:]) <= 1e-4 * cabs)
