"""Oracle solvers (FP64 restatements of PBCGS2, QMRCCG, PETRKP) against a
direct dense solve: the pin for the restated recurrences."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from oracle import solvers as S
from ddscat_hip import targets as T


def small_problem(D=6, m=1.33 + 0.01j, x=1.0):
    tg = T.sphere(D)
    kd = T.kd_for_size_parameter(tg, x)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = np.array([0, 1, 0], complex)
    adia, aoff = T.operators(tg, m, kd, khat, e0)
    b = T.incident_field(tg, khat * kd, e0, tg.centroid_x0())
    A = O.dense_operator(tg.nx, tg.ny, tg.nz, kd, adia.astype(complex), aoff.astype(complex), tg.iocc)
    occ3 = np.flatnonzero(np.tile(tg.iocc != 0, 3))
    Ao = A[np.ix_(occ3, occ3)]
    bo = b.astype(complex)[occ3]
    return Ao, bo


@pytest.mark.parametrize("name", ["pbcgs2", "qmrccg", "petrkp"])
@pytest.mark.parametrize("m", [1.33 + 0.01j, 2.0 + 0.5j])
def test_oracle_solver_converges_to_dense_solution(name, m):
    Ao, bo = small_problem(m=m)
    xref = np.linalg.solve(Ao, bo)
    mv = lambda cw, v: Ao @ v if cw == "N" else np.conj(Ao @ np.conj(v))
    x, it, res, nc = getattr(S, name)(mv, bo, maxit=500, tol=1e-10)
    assert np.linalg.norm(Ao @ x - bo) / np.linalg.norm(bo) < 2e-10
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) < 1e-7
    assert res < 2e-10 and it > 0 and nc >= it


def test_oracle_pbcgs2_nonzero_initial_guess():
    Ao, bo = small_problem()
    xref = np.linalg.solve(Ao, bo)
    mv = lambda cw, v: Ao @ v
    x0 = xref + 0.3 * np.random.default_rng(1).normal(size=xref.size)
    x, it, res, nc = S.pbcgs2(mv, bo, x0=x0, maxit=500, tol=1e-10)
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) < 1e-7


def test_complex_symmetry_makes_cmatvec_adjoint():
    """'C' MATVEC = conj(A conj v) equals A^H v because A^T = A (PETRKP relies on it)."""
    Ao, _ = small_problem(D=4)
    assert np.abs(Ao - Ao.T).max() == 0.0
    v = np.random.default_rng(2).normal(size=Ao.shape[0]) + 1j
    assert np.allclose(np.conj(Ao @ np.conj(v)), Ao.conj().T @ v, rtol=1e-13, atol=1e-13)
