"""ddscat_hip_scat_f32_noalloc (far-field sums with a file-scope device
scratch array, no device allocation) against ddscat_cuda_scat_f32 (per-call
workspace, cu:957-1002) through the C ABI: two calls in a row, the second on
a smaller target.  The per-direction arithmetic is the same kernels in the
same order, only the direction slabs differ, so the results are identical."""
import numpy as np
import pytest

from ddscat_hip import targets as T

pytestmark = pytest.mark.gpu


def solve_sphere(lib, d, xsize, m):
    tg = T.sphere(d)
    kd = T.kd_for_size_parameter(tg, xsize)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = np.array([0, 1, 0], complex)
    adia, aoff = T.operators(tg, m, kd, khat, e0)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    b = T.incident_field(tg, khat * kd, e0, tg.centroid_x0())
    lib.solve("pbcgs2", b, adia, aoff, maxit=200, tol=1e-5)
    return tg, kd, khat, e0


def selected(kd, n):
    th = np.linspace(0.0, np.pi, n)
    aks = np.stack([np.cos(th), np.sin(th), np.zeros(n)], 1) * kd
    em1 = np.stack([-np.sin(th), np.cos(th), np.zeros(n)], 1)
    em2 = np.tile([0.0, 0.0, 1.0], (n, 1))
    return aks, em1, em2


@pytest.mark.parametrize("d,xsize,etasca,nsel", [(32, 2.0, 0.05, 10), (12, 1.0, 0.5, 3)])
def test_noalloc_scat_equals_scat(lib, d, xsize, etasca, nsel):
    tg, kd, khat, e0 = solve_sphere(lib, d, xsize, 1.33 + 0.01j)
    aks, em1, em2 = selected(kd, nsel)
    args = (khat * kd, (1.0, 1.0, 1.0), tg.centroid_x0(), e0, etasca)
    ref = lib.scat(*args, aks=aks, em1=em1, em2=em2)
    got = lib.scat(*args, aks=aks, em1=em1, em2=em2, noalloc=True)
    print(f"D={d}: {ref['navg']} integration directions, Csca {ref['csca']:.10e} vs {got['csca']:.10e}")
    assert got["rc"] == 0 and got["navg"] == ref["navg"] > 0
    for k in ("csca", "cscag2", "cbksca"):
        assert got[k] == ref[k], k
    assert np.array_equal(got["cscag"], ref["cscag"])
    assert np.array_equal(got["f1"], ref["f1"]) and np.array_equal(got["f2"], ref["f2"])
