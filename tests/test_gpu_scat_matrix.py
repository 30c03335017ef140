"""S-matrix / Mueller output path on the device (ddscat_hip_scat_matrix).

Bars: amplitudes f_ij within 1e-5 relative (max norm) of the FP64 oracle on
the same polarizations (the device sums in FP64 and rounds f to complex64);
Mueller elements within 1e-5 of the largest element; the f_1j, f_2j of each
polarization equal to what the reference-ABI SCAT returns for it (same sums,
1e-6); the end-to-end device chain on a sphere within DDA accuracy of Mie
(S11 6 %, normalised elements 0.03, off-block elements < 2 % of S11)."""
import math

import numpy as np
import pytest

from oracle import ddscat_oracle as O
from oracle import scattering as SC
from ddscat_hip import targets as T
from ddscat_hip import scattering as DS

pytestmark = pytest.mark.gpu


def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def pair_frame(rng):
    k = unit(rng.normal(size=3))
    e1 = unit(np.cross(k, rng.normal(size=3)))
    return k, e1.astype(complex), np.cross(k, e1).astype(complex)


def test_scat_matrix_matches_oracle(lib):
    import torch
    tg = T.sphere(10)
    kd = 0.3
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (kd, 0.0, 0.0))
    rng = np.random.default_rng(11)
    npair, n0 = 2, 3 * tg.nat0
    P = (rng.normal(size=(2 * npair, n0)) + 1j * rng.normal(size=(2 * npair, n0))).astype(np.complex64)
    x0 = tg.centroid_x0()
    th = np.radians(np.arange(0, 181, 20.0))
    ph = np.radians(np.arange(0, 360, 45.0))
    TH, PH = (a.reshape(-1) for a in np.meshgrid(th, ph, indexing="ij"))
    ak, e0, aks, em1, em2, frames = [], [], [], [], [], []
    for q in range(npair):
        k, e1, e2 = pair_frame(rng)
        ks, epar, eperp = DS.directions(k, e1, TH, PH)
        ak.append(kd * k); e0.append([e1, e2]); aks.append(kd * ks); em1.append(epar); em2.append(eperp)
        frames.append((k, e1, e2, ks, epar, eperp))
    Pd = torch.from_numpy(P).cuda()
    torch.cuda.synchronize()
    f, M = lib.scat_matrix(Pd.data_ptr(), npair, ak, e0, aks, em1, em2, x0)
    for q, (k, e1, e2, ks, epar, eperp) in enumerate(frames):
        fr = [SC.amplitudes(O.farfield_sums(P[2 * q + j], tg.occ, tg.nx, tg.ny, (1, 1, 1), x0, kd * ks),
                            kd * k, epar, eperp) for j in range(2)]
        fref = np.stack([fr[0][0], fr[0][1], fr[1][0], fr[1][1]], axis=1)
        ferr = np.max(np.abs(f[q] - fref)) / np.max(np.abs(fref))
        Mref = SC.mueller(*SC.amplitude_matrix(fref[:, 0], fref[:, 1], fref[:, 2], fref[:, 3], e1, e2, k, eperp))
        merr = np.max(np.abs(M[q] - Mref)) / np.max(np.abs(Mref))
        print(f"pair {q}: amplitudes {ferr:.2e}, Mueller {merr:.2e}")
        assert ferr < 1e-5
        assert merr < 1e-5
    lib.release()


def test_scat_matrix_amplitudes_equal_reference_scat(lib):
    """f_11, f_21 of the first polarization equal the F1, F2 that the
    reference-ABI SCAT returns after the same solve."""
    import torch
    tg = T.sphere(12)
    m, kd = 1.33 + 0.05j, 0.2
    khat = np.array([1.0, 0.0, 0.0])
    e01 = np.array([0, 1, 0], complex)
    e02 = np.cross(khat, e01).astype(complex)
    x0 = tg.centroid_x0()
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    occ = tg.occ
    comp = lambda v: np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, occ].reshape(-1))
    th = np.radians(np.array([0.0, 30.0, 90.0, 150.0, 180.0]))
    ks, epar, eperp = DS.directions(khat, e01, th, np.full(th.size, 0.4))
    Pc, ref = [], []
    for e0 in (e01, e02):
        adia, aoff = T.operators(tg, m, kd, khat, e0)
        b = T.incident_field(tg, khat * kd, e0, x0)
        x, it, tole, nc, rc = lib.solve("qmrccg", b, adia, aoff, maxit=300, tol=1e-6)
        ref.append(lib.scat(khat * kd, (1, 1, 1), x0, e0, 0.5, aks=kd * ks, em1=epar, em2=eperp))
        Pc.append(comp(x))
    Pd = torch.from_numpy(np.stack(Pc).astype(np.complex64)).cuda()
    torch.cuda.synchronize()
    f, _ = lib.scat_matrix(Pd.data_ptr(), 1, [khat * kd], [[e01, e02]], [kd * ks], [epar], [eperp], x0,
                           mueller=False)
    for j in range(2):
        for i, key in enumerate(("f1", "f2")):
            a, r = f[0][:, 2 * j + i], ref[j][key]
            assert np.max(np.abs(a - r)) <= 1e-6 * np.max(np.abs(r)), (j, key)
    lib.release()


def test_device_chain_sphere_mueller_matches_mie(lib):
    """setup_rhs -> batched PBCGS2 (both polarizations) -> scat_matrix on a
    D=16 sphere, against the Mie series."""
    import torch
    tg = T.sphere(16)
    m, xs = 1.33 + 0.05j, 1.5
    kd = T.kd_for_size_parameter(tg, xs)
    khat = np.array([1.0, 0.0, 0.0])
    e01 = np.array([0, 1, 0], complex)
    e02 = np.cross(khat, e01).astype(complex)
    x0 = tg.centroid_x0()
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    n0 = 3 * tg.nat0
    dev = lambda shape: torch.zeros(shape, dtype=torch.complex64, device="cuda")
    b, x, ad, ao = dev((2, n0)), dev((2, n0)), dev((2, n0)), dev((2, n0))
    eps = torch.full((tg.nat0,), complex(m * m), dtype=torch.complex64, device="cuda")
    lib.setup_rhs([khat * kd, khat * kd], [e01, e02], x0, eps.data_ptr(), b.data_ptr(), ad.data_ptr(),
                  ao.data_ptr())
    it, te, nc, rc = lib.solve_batch("pbcgs2", 2, b.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                     maxit=300, tol=1e-6)
    assert rc == 0
    th = np.radians(np.arange(0, 181, 15.0))
    res = []
    for phi in (0.0, 0.5 * math.pi, 1.1):
        ks, epar, eperp = DS.directions(khat, e01, th, np.full(th.size, phi))
        _, M = lib.scat_matrix(x.data_ptr(), 1, [khat * kd], [[e01, e02]], [kd * ks], [epar], [eperp], x0)
        res.append(M[0])
    m1, m2 = SC.mie_s12(kd * tg.a_eff(), m, th)
    Mm = SC.mueller(m1, m2, 0 * m1, 0 * m1)
    for M in res:
        s11 = M[:, 0, 0]
        print("S11 DDA/Mie:", np.round(s11 / Mm[:, 0, 0], 3))
        assert np.all(np.abs(s11 / Mm[:, 0, 0] - 1) < 0.06)
        for (i, j) in ((0, 1), (2, 2), (2, 3), (3, 3)):
            assert np.all(np.abs(M[:, i, j] / s11 - Mm[:, i, j] / Mm[:, 0, 0]) < 0.03), (i, j)
        for (i, j) in ((0, 2), (0, 3), (1, 2), (1, 3), (2, 0), (3, 0)):
            assert np.all(np.abs(M[:, i, j]) < 0.02 * s11), (i, j)
    avg = DS.orientation_average(np.stack(res))
    assert avg.shape == (th.size, 4, 4)
    lib.release()


def test_scat_matrix_argument_errors(lib):
    tg = T.sphere(6)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (0.3, 0.0, 0.0))
    from ddscat_hip.abi import DDSCATError
    with pytest.raises(DDSCATError) as e:
        lib.scat_matrix(0, 1, [[0.3, 0, 0]], [[[0, 1, 0], [0, 0, 1]]], [[[0.3, 0, 0]]], [[[0, 0, 1]]],
                        [[[0, -1, 0]]], (0, 0, 0))
    assert e.value.code == 340
    lib.release()
