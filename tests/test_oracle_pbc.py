"""Pins of the IPBC=1 (periodic target) restatement in the oracle
(cu:634-645 replica sums, cu:650 whole-grid transform, cu:113 whole-grid
Green multiply).

Nothing in the reference's files holds PBC vectors, so the oracle is pinned by
what the mathematics fixes:
  * Bloch periodicity: the replica sum obeys G(l + P) = e^{i k.P} G(l) for a
    shift by one period P (the damping window moves with the lag), which fixes
    the replica spacing and the sign of the Bloch phase;
  * the replica window against a brute-force sum over 40 more replicas on
    every side with the same damping, which fixes the window bounds;
  * with k along x and a period far beyond the damping range, the lags on the
    stationary-phase line (ly = 0) get exactly the isolated Green tensor,
    which fixes krf / phasf (any error there damps the direct term);
  * the damped sums converge as GAMMA -> 0 (differences shrink), which fixes
    the sign of the damping;
  * the FFT MATVEC with the whole spectral grid equals the dense operator
    built straight from the replica sums at r_j - r_l; the PBC lag function is
    not even, so this also fixes the lag orientation and shows that the
    octant reconstruction of IPBC=0 would be wrong here."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O

AK = (0.3, 0.12, -0.07)
CASES = [(5.0, 0.0), (0.0, 4.0), (5.0, 4.0)]  # y-periodic, z-periodic, both


@pytest.mark.parametrize("pyd,pzd", CASES)
def test_bloch_periodicity(pyd, pzd):
    for lag in ((2, -1, 1), (0, 2, -3), (-1, 0, 0)):
        g0 = O.green_sum_lag(*lag, AK, 0.05, pyd, pzd)
        if pyd > 0:
            g1 = O.green_sum_lag(lag[0], lag[1] + int(pyd), lag[2], AK, 0.05, pyd, pzd)
            assert np.max(np.abs(g1 - np.exp(1j * AK[1] * pyd) * g0)) < 1e-12 * np.max(np.abs(g0))
        if pzd > 0:
            g1 = O.green_sum_lag(lag[0], lag[1], lag[2] + int(pzd), AK, 0.05, pyd, pzd)
            assert np.max(np.abs(g1 - np.exp(1j * AK[2] * pzd) * g0)) < 1e-12 * np.max(np.abs(g0))


@pytest.mark.parametrize("pyd,pzd", CASES)
def test_window_holds_every_significant_replica(pyd, pzd):
    lag = (2, -1, 1)
    g = O.green_sum_lag(*lag, AK, 0.05, pyd, pzd)
    jpbc, jpy1, jpy2, jpz1, jpz2, _, _ = O.pbc_window(lag[0], lag[1], lag[2], AK, 0.05, pyd, pzd)
    wide = (jpy1 - 40, jpy2 + 40) if jpbc != 2 else (0, 0)
    wide = wide + ((jpz1 - 40, jpz2 + 40) if jpbc != 1 else (0, 0))
    gb = O.green_sum_lag(*lag, AK, 0.05, pyd, pzd, window=wide)
    assert np.max(np.abs(gb - g)) < 1e-6 * np.max(np.abs(g))
    # and the window is not trivially wide: dropping its end replicas changes the sum
    narrow = (jpy1 + 3, jpy2 - 3) if jpbc != 2 else (0, 0)
    narrow = narrow + ((jpz1 + 3, jpz2 - 3) if jpbc != 1 else (0, 0))
    gn = O.green_sum_lag(*lag, AK, 0.05, pyd, pzd, window=narrow)
    assert np.max(np.abs(gn - g)) > 1e-4 * np.max(np.abs(g))


def test_stationary_line_gets_the_isolated_term():
    ak = (0.4, 0.0, 0.0)
    for lag in ((3, 0, 2), (1, 0, -2), (-2, 0, 1)):
        g = O.green_sum_lag(*lag, ak, 0.05, 1.0e4, 0.0)
        gi = O.green_elements(*lag, 0.4)
        assert np.max(np.abs(g - gi)) <= 1e-14 * np.max(np.abs(gi))


def test_damped_sums_converge_as_gamma_shrinks():
    lag = (2, 1, 0)
    g = [O.green_sum_lag(*lag, AK, gm, 3.0, 0.0) for gm in (0.08, 0.04, 0.02, 0.01)]
    d = [np.max(np.abs(g[i + 1] - g[i])) for i in range(3)]
    assert d[2] < d[1] < d[0]
    assert d[2] < 0.5 * d[0]


def test_kperp_zero_is_refused():
    with pytest.raises(ValueError):
        O.pbc_window(1.0, 0.0, 0.0, (0.0, 0.0, 0.3), 0.05, 0.0, 4.0)  # z-periodic, k along z


@pytest.mark.parametrize("pyd,pzd", CASES)
def test_fft_matvec_equals_dense_replica_operator(pyd, pzd):
    nx, ny, nz = 3, 2, 2
    nat = nx * ny * nz
    rng = np.random.default_rng(7)
    iocc = np.ones(nat, np.int16)
    iocc[[1, 7]] = 0
    gfull = O.green_full_pbc(nx, ny, nz, AK, 0.05, pyd, pzd)
    gm = O.dense_green_matrix_pbc(nx, ny, nz, AK, 0.05, pyd, pzd)
    occ3 = np.tile(iocc != 0, 3)
    x = (rng.normal(size=3 * nat) + 1j * rng.normal(size=3 * nat)) * occ3
    adia = rng.normal(size=3 * nat) + 1j * rng.normal(size=3 * nat)
    aoff = 0.1 * (rng.normal(size=3 * nat) + 1j * rng.normal(size=3 * nat))
    A = O.dense_operator(nx, ny, nz, np.linalg.norm(AK), adia, aoff, iocc, gm=gm)
    y = O.matvec_fft("N", x, adia, aoff, gfull, nx, ny, nz, iocc, ipbc=1)
    assert np.linalg.norm(y - A @ x) < 1e-12 * np.linalg.norm(A @ x)
    yc = O.matvec_fft("C", x, adia, aoff, gfull, nx, ny, nz, iocc, ipbc=1)
    assert np.linalg.norm(yc - np.conj(A @ np.conj(x))) < 1e-12 * np.linalg.norm(yc)
    # the lag orientation matters: the transposed interaction is another operator
    At = O.dense_operator(nx, ny, nz, np.linalg.norm(AK), adia, aoff, iocc,
                          gm=_transposed_lags(gm, nat))
    assert np.linalg.norm(y - At @ x) > 1e-3 * np.linalg.norm(y)


def _transposed_lags(gm, nat):
    """Interaction with G_pbc(r_l - r_j) in place of G_pbc(r_j - r_l): swap the
    site indices inside every 3x3 component block."""
    out = np.empty_like(gm)
    for p in range(3):
        for q in range(3):
            out[p * nat:(p + 1) * nat, q * nat:(q + 1) * nat] = gm[p * nat:(p + 1) * nat, q * nat:(q + 1) * nat].T
    return out


def test_pbc_spectrum_has_no_octant_symmetry():
    """The whole grid is needed: rebuilding it from its octant with the IPBC=0
    parity rules gives a different tensor."""
    nx, ny, nz = 3, 2, 2
    gfull = O.green_full_pbc(nx, ny, nz, AK, 0.05, 5.0, 0.0)
    oct_ = np.ascontiguousarray(gfull[:, :nz + 1, :ny + 1, :nx + 1])
    rebuilt = O.green_full_from_octant(oct_, nx, ny, nz)
    assert np.max(np.abs(rebuilt - gfull)) > 1e-2 * np.max(np.abs(gfull))
