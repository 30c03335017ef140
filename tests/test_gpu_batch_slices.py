"""Batched MATVEC and batched QMRCCG / PBCGS2 on the SLICES (low-memory)
backend, which runs the right-hand sides of a batch one after another through
its slice buffers, every kernel exiting when its RHS's active flag is 0.

Bars: batched SLICES MATVEC bit-identical to the single-RHS SLICES MATVEC of
each RHS (same kernels, same order); batched solve per RHS: true residual
under the FP64 oracle operator (complex128 FFT MATVEC, cu:104-118) <= 2 tol,
solution within 60 tol of the FFT3D backend's batched solve of the same RHS
(each is within 30 tol of the exact solution), iteration count within
max(2, 25 %) of it; graph replay bit-identical to eager launches; a
mixed-backend multi-target call bit-identical to one target after another.
Sphere D=8 runs the LDS-staged SLICES kernels (padded axes of 16 points),
D=16 the register ones (32 points)."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from ddscat_hip import BACKEND_FFT3D, BACKEND_SLICES, targets as T

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()


def host(t):
    return t.cpu().numpy().view(np.complex64)


def rhs_set(tg, m, kd, dirs):
    occ = tg.occ
    comp = lambda v: np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)).astype(np.complex64)
    bs, ads, aos = [], [], []
    x0 = tg.centroid_x0()
    for khat in dirs:
        khat = np.asarray(khat, float) / np.linalg.norm(khat)
        for e0 in T.polarization_pair(khat):
            adia, aoff = T.operators(tg, m, kd, khat, e0)
            bs.append(comp(T.incident_field(tg, khat * kd, e0, x0)))
            ads.append(comp(adia))
            aos.append(comp(aoff))
    return np.stack(bs), np.stack(ads), np.stack(aos)


def make_case(d, m2=None):
    """Two incidences x two polarizations; with m2 the second incidence's
    RHS see a second material (their solves take more iterations, so RHS
    retire at different iterations of the batch)."""
    tg = T.sphere(d)
    kd = T.kd_for_size_parameter(tg, 1.0)
    b, ad, ao = rhs_set(tg, 1.33 + 0.01j, kd, [(1, 0, 0), (0.3, 0.4, 0.866)])
    if m2 is not None:
        b2, ad2, ao2 = rhs_set(tg, m2, kd, [(0.3, 0.4, 0.866)])
        b[2:], ad[2:], ao[2:] = b2, ad2, ao2
    return dict(tg=tg, kd=kd, b=b, ad=ad, ao=ao)


def prepare(lib, case, backend):
    tg = case["tg"]
    lib.release()
    lib.set_backend(backend)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, np.array([1.0, 0.0, 0.0]) * case["kd"])
    assert lib.info()["backend"] == backend
    assert lib.nat0 * 3 == case["b"].shape[1]


@pytest.fixture()
def restore(lib):
    yield lib
    lib.release()
    lib.set_backend(BACKEND_FFT3D)


def solve(lib, case, method, graph):
    import torch
    B = case["b"].shape[0]
    bd, ad, ao = dev(case["b"]), dev(case["ad"]), dev(case["ao"])
    x = torch.zeros_like(bd)
    torch.cuda.synchronize()
    it, te, nc, rc = lib.solve_batch(method, B, bd.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                     op_per_rhs=True, maxit=300, tol=TOL, graph=graph)
    assert rc == 0
    return host(x).reshape(B, -1).copy(), it.copy(), te.copy(), nc.copy()


def oracle_residual(case, r, x):
    tg, kd = case["tg"], case["kd"]
    occ = tg.occ
    goct = O.green_octant(tg.nx, tg.ny, tg.nz, float(np.float32(kd)))
    full = lambda v: O.scatter_compact(np.asarray(v, np.complex128), occ, tg.nat)
    comp = lambda v: np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)
    ax = comp(O.matvec_fft("N", full(x), full(case["ad"][r]), full(case["ao"][r]), goct, tg.nx, tg.ny, tg.nz,
                           tg.iocc))
    b = case["b"][r].astype(np.complex128)
    return float(np.linalg.norm(ax - b) / np.linalg.norm(b))


@pytest.mark.parametrize("d", [8, 16])
def test_slices_apply_batch_bit_identical_to_single(restore, d):
    import torch
    lib = restore
    case = make_case(d)
    prepare(lib, case, BACKEND_SLICES)
    B = case["b"].shape[0]
    rng = np.random.default_rng(d)
    x = (rng.normal(size=case["b"].shape) + 1j * rng.normal(size=case["b"].shape)).astype(np.complex64)
    xd, ad, ao = dev(x), dev(case["ad"]), dev(case["ao"])
    for cw in ("N", "C"):
        yd = torch.empty_like(xd)
        torch.cuda.synchronize()
        lib.apply_batch(cw, B, xd.data_ptr(), yd.data_ptr(), ad.data_ptr(), ao.data_ptr(), op_per_rhs=True)
        lib.sync()
        yb = host(yd).reshape(B, -1)
        n = x.shape[1]
        for r in range(B):
            xs, ys = dev(x[r]), torch.empty(2 * n, dtype=torch.float32, device="cuda")
            a1, a2 = dev(case["ad"][r]), dev(case["ao"][r])
            torch.cuda.synchronize()
            lib.apply_device(cw, xs.data_ptr(), ys.data_ptr(), a1.data_ptr(), a2.data_ptr(), layout=1)
            lib.sync()
            assert np.array_equal(host(ys), yb[r]), f"D={d} {cw} RHS {r}"


@pytest.mark.parametrize("d,m2", [(8, None), (16, None), (16, 2.5 + 0.3j)])
@pytest.mark.parametrize("method", ["qmrccg", "pbcgs2"])
def test_slices_solve_batch_matches_oracle_and_fft3d(restore, d, m2, method):
    lib = restore
    case = make_case(d, m2)
    prepare(lib, case, BACKEND_FFT3D)
    xf, itf, _, _ = solve(lib, case, method, True)
    prepare(lib, case, BACKEND_SLICES)
    runs = [solve(lib, case, method, g) for g in (True, False)]
    xs, its, tes, ncs = runs[0]
    # graph replay and eager launches: bit for bit
    assert np.array_equal(runs[1][0], xs) and np.array_equal(runs[1][1], its)
    for r in range(xs.shape[0]):
        true_res = oracle_residual(case, r, xs[r])
        diff = float(np.linalg.norm(xs[r].astype(complex) - xf[r]) / np.linalg.norm(xf[r]))
        print(f"D={d} {method} RHS {r}: SLICES it={its[r]} nc={ncs[r]} tole={tes[r]:.2e} oracle resid "
              f"{true_res:.2e} | FFT3D it={itf[r]} | x diff {diff:.2e}")
        assert tes[r] <= 2 * TOL
        assert true_res <= 2 * TOL
        assert diff <= 60 * TOL
        assert abs(int(its[r]) - int(itf[r])) <= max(2, 0.25 * itf[r])


def test_slices_apply_batch_pbc_bit_identical_to_single(restore):
    """IPBC=1 (periodic in y and z, whole-grid Green, register y pass at
    2 NY = 32) on SLICES: the batched MATVEC of 3 RHS equals each RHS's
    single MATVEC bit for bit."""
    import torch
    lib = restore
    nx, ny, nz = 4, 16, 3
    rng = np.random.default_rng(1643)
    iocc = (rng.random(nx * ny * nz) < 0.7).astype(np.int16)
    iocc[0] = 1
    n0 = 3 * int(np.count_nonzero(iocc))
    B = 3
    cx = lambda *shape: (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    x, ad, ao = cx(B, n0), cx(B, n0) + np.complex64(3.0), cx(B, n0) * np.complex64(0.1)
    lib.release()
    lib.set_backend(BACKEND_SLICES)
    lib.prepare_green(nx, ny, nz, iocc, (np.float32(0.3), np.float32(0.12), np.float32(-0.07)), ipbc=1,
                      gamma=np.float32(0.05), pyd=16.0, pzd=3.0)
    assert lib.nat0 * 3 == n0
    xd, add, aod = dev(x), dev(ad), dev(ao)
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    lib.apply_batch("N", B, xd.data_ptr(), yd.data_ptr(), add.data_ptr(), aod.data_ptr(), op_per_rhs=True)
    lib.sync()
    yb = host(yd).reshape(B, -1)
    for r in range(B):
        xs, ys = dev(x[r]), torch.empty(2 * n0, dtype=torch.float32, device="cuda")
        a1, a2 = dev(ad[r]), dev(ao[r])
        torch.cuda.synchronize()
        lib.apply_device("N", xs.data_ptr(), ys.data_ptr(), a1.data_ptr(), a2.data_ptr(), layout=1)
        lib.sync()
        assert np.array_equal(host(ys), yb[r]), f"RHS {r}"


def test_mixed_backend_targets_match_sequential(restore):
    """Slot 1 on SLICES (D=16) and slot 2 on FFT3D (D=8) solved in one
    ddscat_hip_solve_targets call equal their one-at-a-time batched solves."""
    import torch
    lib = restore
    cases = [make_case(16, 2.5 + 0.3j), make_case(8)]
    backends = [BACKEND_SLICES, BACKEND_FFT3D]
    bufs = []
    try:
        for t, (case, be) in enumerate(zip(cases, backends)):
            lib.select_target(t + 1)
            prepare(lib, case, be)
            bufs.append(dict(b=dev(case["b"]), ad=dev(case["ad"]), ao=dev(case["ao"]), B=case["b"].shape[0]))
        seq = []
        for t, d in enumerate(bufs):
            lib.select_target(t + 1)
            x = torch.zeros_like(d["b"])
            torch.cuda.synchronize()
            lib.solve_batch("pbcgs2", d["B"], d["b"].data_ptr(), x.data_ptr(), d["ad"].data_ptr(),
                            d["ao"].data_ptr(), maxit=300, tol=TOL)
            seq.append(x)
        xs = [torch.zeros_like(d["b"]) for d in bufs]
        torch.cuda.synchronize()
        per, rc = lib.solve_targets("pbcgs2", [1, 2], [d["B"] for d in bufs], [d["b"].data_ptr() for d in bufs],
                                    [x.data_ptr() for x in xs], [d["ad"].data_ptr() for d in bufs],
                                    [d["ao"].data_ptr() for d in bufs], maxit=300, tol=TOL)
        torch.cuda.synchronize()
        assert rc == 0
        for t in range(2):
            assert np.all(per[t][1] <= TOL), per[t][1]
            assert torch.equal(xs[t], seq[t]), f"slot {t + 1}: concurrent solve differs from the sequential one"
    finally:
        for t in range(2):
            lib.select_target(t + 1)
            lib.release()
            lib.set_backend(BACKEND_FFT3D)
        lib.select_target(0)
