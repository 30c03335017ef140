"""Scattering directions and polarization bases for the S-matrix / Mueller
output path (``Library.scat_matrix``, ``ddscat_hip_scat_matrix``).

Conventions (Bohren & Huffman, ch. 3): for incidence along khat and a
scattering direction at polar angle theta and azimuth phi, with phi measured
from a1 = Re(e01) (made orthogonal to khat) toward a2 = khat x a1, which is
the frame the reference's SCAT quadrature uses (ddscat_matvec_cuda.cu:849-875):
    khat_s = cos(theta) khat + sin(theta) (cos(phi) a1 + sin(phi) a2)
    e_perp = sin(phi) a1 - cos(phi) a2          (normal to the scattering plane)
    e_par_s = khat_s x e_perp,  e_par_i = khat x e_perp
so e_perp is defined at theta = 0 and pi as well.
"""
from __future__ import annotations

import numpy as np


def directions(khat, e01, theta, phi):
    """(khat_s, e_par_s, e_perp), each [n, 3], for angles theta, phi (radians)."""
    k = np.asarray(khat, np.float64)
    k = k / np.linalg.norm(k)
    a1 = np.real(np.asarray(e01, np.complex128))
    a1 = a1 - np.dot(a1, k) * k
    nrm = np.linalg.norm(a1)
    if not nrm > 0:
        raise ValueError("Re(e01) must have a component across khat")
    a1 = a1 / nrm
    a2 = np.cross(k, a1)
    t = np.asarray(theta, np.float64).reshape(-1)[:, None]
    p = np.broadcast_to(np.asarray(phi, np.float64).reshape(-1), t.shape[:1])[:, None]
    ks = np.cos(t) * k[None] + np.sin(t) * (np.cos(p) * a1[None] + np.sin(p) * a2[None])
    e_perp = np.sin(p) * a1[None] - np.cos(p) * a2[None]
    return ks, np.cross(ks, e_perp), e_perp


def orientation_average(mueller, weights=None):
    """Weighted mean over the leading (orientation) axis of Mueller matrices
    [npair, ndir, 4, 4]; equal weights by default."""
    m = np.asarray(mueller, np.float64)
    w = np.ones(m.shape[0]) if weights is None else np.asarray(weights, np.float64)
    return np.tensordot(w / w.sum(), m, axes=(0, 0))
