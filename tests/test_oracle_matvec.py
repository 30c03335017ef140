"""Pins of the oracle MATVEC (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:104-118,
ddscat_cuda_unified/fortran/cprod_cuda.f90:446-607) and of the committed golden fixtures."""
from pathlib import Path

import numpy as np
import pytest

from oracle import ddscat_oracle as O

GOLDEN = Path(__file__).parent / "golden"


def case(rng, nx, ny, nz, fill=0.7):
    nat = nx * ny * nz
    iocc = (rng.random(nat) < fill).astype(np.int16)
    iocc[0] = 1
    cx = lambda s=1.0: s * (rng.normal(size=3 * nat) + 1j * rng.normal(size=3 * nat))
    return iocc, cx() * np.tile(iocc != 0, 3), cx(), cx(0.3)


@pytest.mark.parametrize("dims", [(1, 1, 2), (2, 2, 2), (3, 4, 5), (5, 3, 4), (6, 2, 3)])
@pytest.mark.parametrize("akd", [0.05, 0.6])
def test_fft_matvec_equals_dense_operator(dims, akd):
    """Brute force: dense 3NAT x 3NAT matrix from the closed form, no FFT."""
    rng = np.random.default_rng(hash(dims) % 1000)
    iocc, x, adia, aoff = case(rng, *dims)
    go = O.green_octant(*dims, akd)
    y = O.matvec_fft("N", x, adia, aoff, go, *dims, iocc)
    A = O.dense_operator(*dims, akd, adia, aoff, iocc)
    assert np.linalg.norm(y - A @ x) <= 1e-13 * np.linalg.norm(A @ x)


def test_conjugate_mode_identity_and_complex_symmetry():
    rng = np.random.default_rng(4)
    dims = (4, 3, 5)
    iocc, x, adia, aoff = case(rng, *dims)
    _, y, *_ = case(rng, *dims)
    go = O.green_octant(*dims, 0.33)
    yc = O.matvec_fft("C", x, adia, aoff, go, *dims, iocc)
    assert np.allclose(yc, np.conj(O.matvec_fft("N", np.conj(x), adia, aoff, go, *dims, iocc)), rtol=1e-13)
    # x^T A y = y^T A x (A = A^T for complex-symmetric polarizability blocks)
    m = np.tile(iocc != 0, 3)
    x, y = x * m, y * m
    Ay = O.matvec_fft("N", y, adia, aoff, go, *dims, iocc)
    Ax = O.matvec_fft("N", x, adia, aoff, go, *dims, iocc)
    assert abs(np.dot(x, Ay) - np.dot(y, Ax)) < 1e-12 * abs(np.dot(x, Ay))


def test_offdiagonal_block_ordering():
    """aoff = [a23, a31, a12]: the local 3x3 block is symmetric with those entries (cp:503-515)."""
    dims = (1, 1, 1)
    go = np.zeros((6, 2, 2, 2), complex)
    adia = np.array([1.0, 2.0, 3.0], complex)
    aoff = np.array([10.0, 20.0, 30.0], complex)  # a23, a31, a12
    x = np.array([1.0, 100.0, 10000.0], complex)
    y = O.matvec_fft("N", x, adia, aoff, go, *dims)
    M = np.array([[1, 30, 20], [30, 2, 10], [20, 10, 3]], complex)
    assert np.allclose(y, M @ x)


def test_compact_mapping_round_trip():
    """Restates the reference's layout self-check (scripts/check_compact_mapping.py:3-20)."""
    nat = 10
    occ = np.array([0, 2, 5, 9])
    full = (np.arange(3 * nat) + 1j * np.arange(3 * nat) / 10).astype(np.complex64)
    comp = O.gather_compact(full, occ, nat)
    assert comp.size == 3 * occ.size
    for q in range(3 * occ.size):
        c, k = divmod(q, occ.size)
        assert comp[q] == full[c * nat + occ[k]]
    back = O.scatter_compact(comp, occ, nat)
    expect = np.zeros_like(full)
    for c in range(3):
        expect[c * nat + occ] = full[c * nat + occ]
    assert np.array_equal(back, expect)


def test_cpu_fp32_baseline_operator_matches_oracle():
    rng = np.random.default_rng(8)
    dims = (6, 5, 4)
    iocc, x, adia, aoff = case(rng, *dims)
    go = O.green_octant(*dims, 0.2)
    ref = O.matvec_fft("N", x, adia, aoff, go, *dims, iocc)
    got = O.CpuMatvecFP32(go, *dims, workers=2).apply("N", x, adia, aoff, iocc)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


def test_golden_fixtures_reproduce():
    """Committed vectors (tests/golden/make_golden.py) still come out of the oracle."""
    f = np.load(GOLDEN / "matvec_small.npz", allow_pickle=False)
    dims = tuple(int(v) for v in f["dims"])
    akd = float(f["akd"])
    go = O.green_octant(*dims, akd)
    assert np.allclose(go, f["green_octant"], rtol=1e-12, atol=1e-12)
    for cw in ("N", "C"):
        y = O.matvec_fft(cw, f["x"], f["adia"], f["aoff"], go, *dims, f["iocc"])
        assert np.allclose(y, f[f"y_{cw}"], rtol=1e-12, atol=1e-12)
    A = O.dense_operator(*dims, akd, f["adia"], f["aoff"], f["iocc"])
    assert np.allclose(A @ f["x"], f["y_N"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dims", [(3, 4, 5), (6, 2, 3), (8, 7, 6)])
def test_pruned_fp64_operator_equals_matvec_fft(dims):
    """FftMatvec64 (the large-grid oracle operator: pruned transforms on
    compact vectors) is matvec_fft, in both modes, on a grid with holes."""
    rng = np.random.default_rng(sum(dims))
    iocc, x, adia, aoff = case(rng, *dims)
    go = O.green_octant(*dims, 0.33)
    op = O.FftMatvec64(go, *dims, iocc, workers=2)
    occ = O.occupied_index(iocc)
    nat = int(np.prod(dims))
    comp = lambda v: O.gather_compact(v, occ, nat)
    for cw in ("N", "C"):
        ref = comp(O.matvec_fft(cw, x, adia, aoff, go, *dims, iocc))
        got = op(cw, comp(x), comp(adia), comp(aoff))
        assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)
