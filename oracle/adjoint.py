"""FP64 CPU adjoint of the absorption cross section (test infrastructure).

Only tests/ and bench.py's CPU leg use this module, as the checker of the
device adjoint (ddscat_hip_cabs_adjoint); no product path imports it.

Cabs is the EVALQ quantity (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:
939-950): Cabs = -(4 pi k / E0^2) Im(P^H M P), with M the per-site matrix
diag(adia + i (2/3) k^3) plus the symmetric off-diagonal aoff = [a23, a31, a12]
(cu:770-783), and P = A^{-1} E.  Writing W = -(4 pi k / E0^2) Im(M) (real
symmetric per site), Cabs = P^H W P and

  dCabs = P^H dW P + 2 Re(P^H W dP),  dP = -A^{-1} dA P.

A is complex symmetric, so lam = A^{-H} W P is conj(mu) with A mu = W conj(P),
and 2 Re(P^H W dP) = -2 Re(mu^T dA P).  The returned g_adia, g_aoff satisfy
dCabs = sum Re(g * d(operator entry)).  Pinned by central finite differences
of cabs_dense (tests/test_oracle_adjoint.py)."""
from __future__ import annotations

import numpy as np

PI = np.pi


def site_matrices(adia_c, aoff_c, rabs=0.0):
    """Per-site 3x3 matrices of the operator's local part: [nat0, 3, 3]."""
    ad = np.asarray(adia_c, np.complex128).reshape(3, -1)
    a23, a31, a12 = np.asarray(aoff_c, np.complex128).reshape(3, -1)
    n = ad.shape[1]
    M = np.zeros((n, 3, 3), np.complex128)
    for c in range(3):
        M[:, c, c] = ad[c] + 1j * rabs
    M[:, 0, 1] = M[:, 1, 0] = a12
    M[:, 0, 2] = M[:, 2, 0] = a31
    M[:, 1, 2] = M[:, 2, 1] = a23
    return M


def cabs_of(p_c, adia_c, aoff_c, akk, e02=1.0):
    """Cabs of a compact polarization [3][NAT0] (FP64 restatement of EVALQ's Cabs)."""
    p = np.asarray(p_c, np.complex128).reshape(3, -1).T  # [nat0, 3]
    M = site_matrices(adia_c, aoff_c, akk ** 3 / 1.5)
    q = np.einsum("nij,nj->ni", M, p)
    return -4.0 * PI * akk * np.vdot(p.reshape(-1), q.reshape(-1)).imag / e02


def cabs_dense(Ao, e_c, adia_c, aoff_c, akk, e02=1.0):
    """Forward model on the dense compact operator: P = Ao^{-1} E, then Cabs."""
    p = np.linalg.solve(Ao, np.asarray(e_c, np.complex128))
    return cabs_of(p, adia_c, aoff_c, akk, e02), p


def adjoint_rhs(p_c, adia_c, aoff_c, akk, e02=1.0):
    """W conj(P) [3*NAT0]: the right-hand side of the adjoint solve A mu = W conj(P)."""
    n0 = np.asarray(p_c).size // 3
    p = np.asarray(p_c, np.complex128).reshape(3, n0)
    cf = 4.0 * PI * akk / e02
    W = -cf * site_matrices(adia_c, aoff_c, akk ** 3 / 1.5).imag  # real symmetric per site
    return np.einsum("nij,nj->ni", W, np.conj(p.T)).T.reshape(-1)


def gradient_terms(p_c, mu_c, akk, e02=1.0):
    """g_adia, g_aoff [3*NAT0] from the solution p and the adjoint solution mu."""
    n0 = np.asarray(p_c).size // 3
    p = np.asarray(p_c, np.complex128).reshape(3, n0)
    mu = np.asarray(mu_c, np.complex128).reshape(3, n0)
    cf = 4.0 * PI * akk / e02
    g_adia = 1j * cf * np.abs(p) ** 2 - 2.0 * mu * p
    g_aoff = np.empty_like(g_adia)
    for s, (a, b) in enumerate(((1, 2), (2, 0), (0, 1))):  # a23, a31, a12
        g_aoff[s] = 1j * cf * 2.0 * np.real(np.conj(p[a]) * p[b]) - 2.0 * (mu[a] * p[b] + mu[b] * p[a])
    return g_adia.reshape(-1), g_aoff.reshape(-1)


def cabs_adjoint(Ao, p_c, adia_c, aoff_c, akk, e02=1.0):
    """g_adia, g_aoff [3][NAT0] with dCabs = sum Re(g . d entry), for the
    solution p of Ao p = E.  Ao: the dense compact operator, or a callable
    solving Ao mu = rhs (the iterative oracle solve on large grids)."""
    rhs = adjoint_rhs(p_c, adia_c, aoff_c, akk, e02)
    mu = Ao(rhs) if callable(Ao) else np.linalg.solve(Ao, rhs)
    return gradient_terms(p_c, mu, akk, e02)


def compact_dense_operator(dense_operator, nx, ny, nz, akd, adia_c, aoff_c, iocc):
    """Dense compact operator A[occ3, occ3] with compact operators scattered to full."""
    iocc = np.asarray(iocc)
    occ = np.flatnonzero(iocc)
    nat = nx * ny * nz
    ad = np.zeros((3, nat), complex)
    ao = np.zeros((3, nat), complex)
    ad[:, occ] = np.asarray(adia_c).reshape(3, -1)
    ao[:, occ] = np.asarray(aoff_c).reshape(3, -1)
    A = dense_operator(nx, ny, nz, akd, ad.reshape(-1), ao.reshape(-1), iocc)
    occ3 = np.flatnonzero(np.tile(iocc != 0, 3))
    return A[np.ix_(occ3, occ3)]
