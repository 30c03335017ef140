"""Deterministic inputs of the full-size configuration tests.

Shared by the golden-fixture script (tests/golden/make_config_golden.py), which
solves them with the FP64 oracle in this container, and by the GPU tests that
check the HIP path against those fixtures.  Every input is built from a seed
or a closed form, so both sides regenerate the same arrays; a checksum of the
generated inputs is stored with each fixture and compared before use.

  block128   128^3 block, m = 1.5+0.1i, x = 10, +x incidence, E || y and z
  ellipsoids the first targets of the seeded uneven batch: unions of 1-4
             ellipsoids (semi-axes U[6,40]), m per target, incidences from
             random SO(3) rotations, kd = 0.1
  design64   64^3 block with a smooth random permittivity in [1,4] + 0.05i,
             +z incidence, x = 6, J = mean Q_abs of both polarizations
  sphere64   sphere D=64 with the large-target material (m = 1.7+0.05i,
             kd = 0.25, |m| kd = 0.43), +x incidence, E || y and z
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

from ddscat_hip import targets as T

SEED = 20261202


def checksum(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()[:16]


def compact(tg, v):
    """Full [3*NAT] -> compact [3*NAT0] (IOCC order)."""
    return np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, tg.occ].reshape(-1))


def incident_compact(tg, kvec, e0):
    return compact(tg, T.incident_field(tg, kvec, e0, tg.centroid_x0()))


def geometric_cross_section(tg) -> float:
    """pi a_eff^2 in units of d^2 (Q = C / (pi a_eff^2))."""
    return math.pi * tg.a_eff() ** 2


# ---------------------------------------------------------------- block128 --
def block128(edge=128, m=1.5 + 0.1j, xsize=10.0):
    """The mixed-precision single-target workload: returns dict(tg, kd, khat,
    pols, rhs=[(b_c, adia_c, aoff_c) per polarization]) with complex64
    compact arrays built on the host (LDR of Draine & Goodman 1993)."""
    tg = T.block(edge)
    kd = T.kd_for_size_parameter(tg, xsize)
    khat = np.array([1.0, 0.0, 0.0])
    pols = [np.array([0, 1, 0], complex), np.array([0, 0, 1], complex)]
    rhs = []
    for e0 in pols:
        adia, aoff = T.operators(tg, m, kd, khat, e0)
        rhs.append((incident_compact(tg, khat * kd, e0), compact(tg, adia), compact(tg, aoff)))
    return dict(tg=tg, kd=kd, khat=khat, pols=pols, m=m, rhs=rhs)


# ---------------------------------------------------------------- sphere64 --
def sphere64(diameter=64, m=1.7 + 0.05j, kd=0.25):
    """A mid-size sphere with the D=384 workload's material and lattice
    spacing (the same slow QMRCCG convergence, at 1/216 of the dipoles)."""
    tg = T.sphere(diameter)
    khat = np.array([1.0, 0.0, 0.0])
    pols = [np.array([0, 1, 0], complex), np.array([0, 0, 1], complex)]
    rhs = []
    for e0 in pols:
        adia, aoff = T.operators(tg, m, kd, khat, e0)
        rhs.append((incident_compact(tg, khat * kd, e0), compact(tg, adia), compact(tg, aoff)))
    return dict(tg=tg, kd=kd, khat=khat, pols=pols, m=m, rhs=rhs)


# -------------------------------------------------------------- ellipsoids --
def ellipsoid_batch(n_targets=48, n_orient=8, kd=0.1, seed=SEED):
    """The seeded uneven batch exactly as the batched benchmark draws it:
    per target (ellipsoid union, m, n_orient rotations)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_targets):
        tg = T.ellipsoid_union(rng)
        m = complex(rng.uniform(1.2, 2.0), rng.uniform(0.0, 0.3))
        out.append((tg, m, T.random_rotations(rng, n_orient)))
    return out


def ellipsoid_cases(n_targets=3, n_orient=2, kd=0.1, seed=SEED):
    """The first n_targets targets of the batch with the first n_orient of
    their orientations x 2 polarizations: list of dict(tg, m, aks [B,3]
    float32, e0s [B,3] complex64, kd)."""
    cases = []
    for tg, m, rots in ellipsoid_batch(n_targets=n_targets, kd=kd, seed=seed):
        aks, e0s = [], []
        for R in rots[:n_orient]:
            khat = R[:, 0]
            for e0 in (R[:, 1], R[:, 2]):
                aks.append(khat * kd)
                e0s.append(e0.astype(np.complex64))
        cases.append(dict(tg=tg, m=m, kd=kd, aks=np.asarray(aks, np.float32), e0s=np.asarray(e0s)))
    return cases


def ellipsoid_rhs_host(case):
    """Host (FP64-formula) inputs of one ellipsoid case, per RHS: incident
    field and LDR 1/alpha on the compact sites, from the float32 k d and
    complex64 polarizations the device setup receives."""
    tg = case["tg"]
    eps = complex(case["m"]) ** 2
    out = []
    for ak, e0 in zip(case["aks"], case["e0s"]):
        ak = ak.astype(np.float64)
        kd = float(np.linalg.norm(ak))
        e0 = e0.astype(np.complex128)
        ainv = 1.0 / T.ldr_alpha(np.full(tg.nat0, eps), kd, ak / kd, e0)
        adia = np.tile(ainv, 3)
        out.append((incident_compact(tg, ak, e0), adia, np.zeros_like(adia)))
    return out


# ---------------------------------------------------------------- design64 --
def design64(edge=64, xsize=6.0, seed=SEED):
    """The inverse-design workload: dict(tg, kd, khat, pols, eps [NAT0]
    complex128, geo).  eps = U[1,4] smoothed by a Gaussian of sigma = 3
    dipoles, + 0.05i."""
    import scipy.ndimage as ndi
    tg = T.block(edge)
    kd = T.kd_for_size_parameter(tg, xsize)
    khat = np.array([0.0, 0.0, 1.0])
    pols = list(T.polarization_pair(khat))
    rng = np.random.default_rng(seed)
    field = ndi.gaussian_filter(rng.uniform(1.0, 4.0, size=(edge, edge, edge)), sigma=3.0)
    eps = field.reshape(-1)[tg.occ] + 0.05j
    return dict(tg=tg, kd=kd, khat=khat, pols=pols, eps=eps, geo=geometric_cross_section(tg))


def ldr_inverse(eps, kd, khat, e0):
    """1/alpha (LDR) and its derivative d(1/alpha)/d eps, complex128."""
    khat = np.asarray(khat, float) / np.linalg.norm(khat)
    ea = np.abs(np.asarray(e0, complex))
    S = float(np.sum((khat * ea / np.linalg.norm(ea)) ** 2))
    eps = np.asarray(eps, np.complex128)
    inv = (4.0 * math.pi / 3.0) * (eps + 2.0) / (eps - 1.0) + \
        (T.LDR_B1 + (T.LDR_B2 + T.LDR_B3 * S) * eps) * kd ** 2 - (2.0 / 3.0) * 1j * kd ** 3
    dinv = -4.0 * math.pi / (eps - 1.0) ** 2 + (T.LDR_B2 + T.LDR_B3 * S) * kd ** 2
    return inv, dinv
