"""Batched MATVEC and batched QMRCCG / PBCGS2 against the single-RHS path
and the dense FP64 solve.

Bars: batched MATVEC bit-identical to the single-RHS MATVEC of each RHS (same
kernels, RHS index in the grid); batched solve per RHS: true residual (FP64,
dense) <= 2 tol, solution within 30 tol of the dense solve, iteration count
within max(2, 25 %) of the single-RHS solver; graph replay bit-identical to
eager launches."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from ddscat_hip import targets as T

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()


def host(t):
    return t.cpu().numpy().view(np.complex64)


def rhs_set(tg, m, kd, dirs):
    """Per RHS: compact b, adia, aoff for incidence khat with both polarizations."""
    occ = tg.occ
    comp = lambda v: np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)).astype(np.complex64)
    bs, ads, aos, meta = [], [], [], []
    x0 = tg.centroid_x0()
    for khat in dirs:
        khat = np.asarray(khat, float) / np.linalg.norm(khat)
        for e0 in T.polarization_pair(khat):
            adia, aoff = T.operators(tg, m, kd, khat, e0)
            b = T.incident_field(tg, khat * kd, e0, x0)
            bs.append(comp(b)); ads.append(comp(adia)); aos.append(comp(aoff))
            meta.append((khat, e0))
    return np.stack(bs), np.stack(ads), np.stack(aos), meta


@pytest.fixture(scope="module")
def case(lib):
    tg = T.sphere(8)
    kd = T.kd_for_size_parameter(tg, 1.0)
    dirs = [(1, 0, 0), (0.3, 0.4, 0.866)]
    b, ad, ao, meta = rhs_set(tg, 1.33 + 0.01j, kd, dirs)
    return dict(tg=tg, kd=kd, b=b, ad=ad, ao=ao, meta=meta)


def prep(lib, case):
    """Prepare the case's target (other tests in the module prepare others);
    the batched entry points index vectors by the prepared NAT0."""
    tg = case["tg"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, np.array([1.0, 0.0, 0.0]) * case["kd"])
    assert lib.nat0 * 3 == case["b"].shape[1]


def test_apply_batch_bit_identical_to_single(lib, case):
    import torch
    prep(lib, case)
    tg = case["tg"]
    B = case["b"].shape[0]
    rng = np.random.default_rng(3)
    x = (rng.normal(size=case["b"].shape) + 1j * rng.normal(size=case["b"].shape)).astype(np.complex64)
    xd, ad, ao = dev(x), dev(case["ad"]), dev(case["ao"])
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    lib.apply_batch("N", B, xd.data_ptr(), yd.data_ptr(), ad.data_ptr(), ao.data_ptr(), op_per_rhs=True)
    lib.sync()
    yb = host(yd).reshape(B, -1)
    n = x.shape[1]
    for r in range(B):
        xs, ys = dev(x[r]), torch.empty(2 * n, dtype=torch.float32, device="cuda")
        a1, a2 = dev(case["ad"][r]), dev(case["ao"][r])
        torch.cuda.synchronize()
        lib.apply_device("N", xs.data_ptr(), ys.data_ptr(), a1.data_ptr(), a2.data_ptr(), layout=1)
        lib.sync()
        assert np.array_equal(host(ys), yb[r]), f"RHS {r}"


@pytest.mark.parametrize("shape", [(16, 12, 20), (24, 20, 18), (32, 32, 32)])
def test_apply_batch_shared_operator_matches_oracle(lib, shape):
    """Generic (48, 40, 36 ...) and register (64) FFT kernels, shared operators."""
    import torch
    nx, ny, nz = shape
    rng = np.random.default_rng(nx + ny + nz)
    iocc = (rng.random(nx * ny * nz) < 0.6).astype(np.int16)
    iocc[0] = 1
    occ = np.flatnonzero(iocc)
    lib.prepare_green(nx, ny, nz, iocc, (0.0, 0.2, 0.1))
    B, n0 = 3, 3 * occ.size
    cx = lambda *s: (rng.normal(size=s) + 1j * rng.normal(size=s)).astype(np.complex64)
    x, ad, ao = cx(B, n0), cx(n0), 0.3 * cx(n0)
    xd, add, aod = dev(x), dev(ad), dev(ao)
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    lib.apply_batch("C", B, xd.data_ptr(), yd.data_ptr(), add.data_ptr(), aod.data_ptr(), op_per_rhs=False)
    lib.sync()
    y = host(yd).reshape(B, -1)
    kd = float(np.linalg.norm(np.float32([0.0, 0.2, 0.1])))
    goct = O.green_octant(nx, ny, nz, kd)
    full = lambda v: np.zeros(3 * nx * ny * nz, complex)
    for r in range(B):
        xf, adf, aof = full(0), full(0), full(0)
        xf.reshape(3, -1)[:, occ] = x[r].reshape(3, -1)
        adf.reshape(3, -1)[:, occ] = ad.reshape(3, -1)
        aof.reshape(3, -1)[:, occ] = ao.reshape(3, -1)
        yr = O.matvec_fft("C", xf, adf, aof, goct, nx, ny, nz, iocc).reshape(3, -1)[:, occ].reshape(-1)
        err = np.linalg.norm(y[r] - yr) / np.linalg.norm(yr)
        assert err < 1e-5, (r, err)


@pytest.mark.parametrize("shape", [(128, 6, 4), (4, 6, 256), (4, 6, 384), (6, 256, 4), (384, 4, 6), (256, 5, 4),
                                   (4, 6, 108), (100, 5, 4)])
def test_apply_batch_register_shapes_bit_identical(lib, shape):
    """Batched MATVEC with the 200-, 216-, 256-, 512- and 768-point register
    kernels (passes 3 and 5 by 2 x 256 and 3 x 256 decimation at 512 and
    768, the RHS index in the grid) bit-identical to the single-RHS MATVEC of
    each RHS."""
    import torch
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    iocc = (rng.random(nx * ny * nz) < 0.7).astype(np.int16)
    iocc[0] = 1
    n0 = 3 * int(np.count_nonzero(iocc))
    lib.prepare_green(nx, ny, nz, iocc, (0.25, 0.1, 0.05))
    B = 3
    cx = lambda *s: (rng.normal(size=s) + 1j * rng.normal(size=s)).astype(np.complex64)
    x, ad, ao = cx(B, n0), cx(B, n0), 0.3 * cx(B, n0)
    xd, add, aod = dev(x), dev(ad), dev(ao)
    yd = torch.empty_like(xd)
    info = lib.info()
    base = {p: info[f"kernel_pass{p}"] for p in (3, 5)}
    checked = 0
    for p in (3, 5):  # every pass-3 and pass-5 kernel, batched and single
        kind = -1
        while lib.set_pass_kernel(p, kind, check=False) == 0:
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
                assert np.array_equal(host(ys), yb[r]), (shape, p, kind, r)
            checked += 1
            kind += 1
        lib.set_pass_kernel(p, base[p])
    print(shape, "pass-3/5 kernels checked:", checked, "tuned", [info[f"kernel_pass{p}"] for p in range(1, 6)])
    lib.release()


def dense_solution(case, r):
    tg, kd = case["tg"], case["kd"]
    occ = tg.occ
    ad = np.zeros(3 * tg.nat, complex); ao = np.zeros_like(ad)
    ad.reshape(3, -1)[:, occ] = case["ad"][r].reshape(3, -1)
    ao.reshape(3, -1)[:, occ] = case["ao"][r].reshape(3, -1)
    A = O.dense_operator(tg.nx, tg.ny, tg.nz, float(np.float32(kd)), ad, ao, tg.iocc)
    occ3 = np.flatnonzero(np.tile(tg.iocc != 0, 3))
    Ao = A[np.ix_(occ3, occ3)]
    b = case["b"][r].astype(complex)
    return Ao, b, np.linalg.solve(Ao, b)


@pytest.mark.parametrize("method,graph", [("qmrccg", True), ("qmrccg", False), ("pbcgs2", True),
                                          ("pbcgs2", False)])
def test_solve_batch_matches_dense_and_single(lib, case, method, graph):
    import torch
    prep(lib, case)
    B = case["b"].shape[0]
    bd, ad, ao = dev(case["b"]), dev(case["ad"]), dev(case["ao"])
    x = torch.zeros_like(bd)
    torch.cuda.synchronize()
    it, te, nc, rc = lib.solve_batch(method, B, bd.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                     op_per_rhs=True, maxit=300, tol=TOL, graph=graph)
    assert rc == 0
    xb = host(x).reshape(B, -1).astype(complex)
    for r in range(B):
        Ao, b, xref = dense_solution(case, r)
        true_res = np.linalg.norm(Ao @ xb[r] - b) / np.linalg.norm(b)
        err = np.linalg.norm(xb[r] - xref) / np.linalg.norm(xref)
        # single-RHS device solve of the same RHS
        xs = torch.zeros(2 * b.size, dtype=torch.float32, device="cuda")
        bs, a1, a2 = dev(case["b"][r]), dev(case["ad"][r]), dev(case["ao"][r])
        torch.cuda.synchronize()
        it1, te1, nc1, rc1 = lib.solve_device(method, bs.data_ptr(), xs.data_ptr(), a1.data_ptr(), a2.data_ptr(),
                                              maxit=300, tol=TOL)
        print(f"RHS {r}: batch it={it[r]} nc={nc[r]} tole={te[r]:.2e} true={true_res:.2e} err={err:.2e} | "
              f"single it={it1} nc={nc1}")
        assert true_res <= 2 * TOL
        assert err <= 30 * TOL
        assert te[r] <= 2 * TOL
        assert abs(int(it[r]) - it1) <= max(2, 0.25 * it1)


@pytest.mark.parametrize("method", ["qmrccg", "pbcgs2"])
def test_solve_batch_graph_equals_eager(lib, case, method):
    import torch
    prep(lib, case)
    B = case["b"].shape[0]
    bd, ad, ao = dev(case["b"]), dev(case["ad"]), dev(case["ao"])
    outs = []
    for graph in (True, False, True):
        x = torch.zeros_like(bd)
        torch.cuda.synchronize()
        it, te, nc, rc = lib.solve_batch(method, B, bd.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                         maxit=300, tol=TOL, graph=graph)
        outs.append((host(x).copy(), it.copy(), nc.copy()))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert np.array_equal(o[1], outs[0][1]) and np.array_equal(o[2], outs[0][2])


@pytest.mark.parametrize("method", ["qmrccg", "pbcgs2"])
def test_solve_batch_ignores_uninitialised_vectors(lib, case, method):
    """A batched solve on vectors that hold NaN bit patterns from earlier use
    gives the same solution, bit for bit, as on clean vectors: every vector
    is written before it is read."""
    import torch
    prep(lib, case)
    B = case["b"].shape[0]
    bd, ad, ao = dev(case["b"]), dev(case["ad"]), dev(case["ao"])
    outs = []
    for poison in (False, True):
        if poison:
            lib.poison_workspace()
        x = torch.zeros_like(bd)
        torch.cuda.synchronize()
        it, te, nc, rc = lib.solve_batch(method, B, bd.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                         maxit=300, tol=TOL)
        assert rc == 0 and np.all(te <= TOL)
        outs.append((host(x).copy(), it.copy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_solve_batch_fp32_mode_converges(lib, case):
    import torch
    prep(lib, case)
    from ddscat_hip import PREC_FP32, PREC_MIXED
    B = case["b"].shape[0]
    bd, ad, ao = dev(case["b"]), dev(case["ad"]), dev(case["ao"])
    x = torch.zeros_like(bd)
    torch.cuda.synchronize()
    lib.set_precision(PREC_FP32)
    try:
        it, te, nc, rc = lib.solve_batch("qmrccg", B, bd.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                         maxit=300, tol=TOL)
    finally:
        lib.set_precision(PREC_MIXED)
    xb = host(x).reshape(B, -1).astype(complex)
    for r in range(B):
        Ao, b, _ = dense_solution(case, r)
        assert np.linalg.norm(Ao @ xb[r] - b) / np.linalg.norm(b) <= 3 * TOL


def test_solve_batch_rejects_bad_arguments(lib, case):
    import torch
    prep(lib, case)
    bd = dev(case["b"])
    x = torch.zeros_like(bd)
    _, _, _, rc = lib.solve_batch("qmrccg", 0, bd.data_ptr(), x.data_ptr(), bd.data_ptr(), bd.data_ptr(),
                                  check=False)
    assert rc == 1
    _, _, _, rc = lib.solve_batch("petrkp", 2, bd.data_ptr(), x.data_ptr(), bd.data_ptr(), bd.data_ptr(),
                                  check=False)
    assert rc == 2


def test_evalq_batch_matches_single_evalq(lib, case):
    """Batched device EVALQ equals the reference-ABI EVALQ (compact host
    arrays) of each RHS; both are FP64 sums of the same complex64 data."""
    import torch
    prep(lib, case)
    B = case["b"].shape[0]
    bd, ad, ao = dev(case["b"]), dev(case["ad"]), dev(case["ao"])
    x = torch.zeros_like(bd)
    torch.cuda.synchronize()
    lib.solve_batch("qmrccg", B, bd.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(), maxit=300, tol=TOL)
    aks = np.stack([m[0] * case["kd"] for m in case["meta"]]).astype(np.float32)
    out = lib.evalq_batch(B, bd.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(), aks)
    xb = host(x).reshape(B, -1)
    for r in range(B):
        cabs, cext, cpha, _ = lib.evalq(aks[r], use_resident=False, e_c=case["b"][r], p_c=xb[r],
                                        adia_c=case["ad"][r], aoff_c=case["ao"][r])
        ref = np.array([cabs, cext, cpha])
        print(f"RHS {r}: batch {out[r]} single {ref}")
        assert np.allclose(out[r], ref, rtol=1e-9, atol=0.0)
        assert cext > 0 and cabs > 0


@pytest.mark.parametrize("prescription", ["LATTDR", "GKDLDR"])
def test_setup_rhs_matches_host_operators(lib, prescription):
    """Device incident fields and LDR operators equal the host harness's
    (targets.incident_field / targets.operators, pinned in test_ldr_cpu.py)
    for random incidences (S != 0), both LDR prescriptions."""
    import torch
    tg = T.sphere(16)
    kd = T.kd_for_size_parameter(tg, 2.0)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (kd, 0.0, 0.0))
    rng = np.random.default_rng(21)
    dirs = rng.normal(size=(3, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    m = 1.6 + 0.05j
    aks, e0s, refs = [], [], []
    occ = tg.occ
    comp = lambda v: np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)
    x0 = tg.centroid_x0()
    for khat in dirs:
        for e0 in T.polarization_pair(khat):
            aks.append(khat * kd); e0s.append(e0)
            ad, _ = T.operators(tg, m, kd, khat, e0, prescription=prescription)
            refs.append((comp(T.incident_field(tg, khat * kd, e0, x0)), comp(ad)))
    B, n = len(aks), 3 * tg.nat0
    eps = torch.full((tg.nat0,), complex(m) ** 2, dtype=torch.complex64, device="cuda")
    b = torch.empty((B, n), dtype=torch.complex64, device="cuda")
    ad = torch.empty_like(b); ao = torch.empty_like(b)
    lib.set_polarizability(prescription)
    try:
        lib.setup_rhs(np.asarray(aks), np.asarray(e0s), x0, eps.data_ptr(), b.data_ptr(), ad.data_ptr(),
                      ao.data_ptr())
    finally:
        lib.set_polarizability("LATTDR")
    b, ad, ao = b.cpu().numpy(), ad.cpu().numpy(), ao.cpu().numpy()
    for r, (bref, adref) in enumerate(refs):
        eb = np.linalg.norm(b[r] - bref) / np.linalg.norm(bref)
        ea = np.linalg.norm(ad[r] - adref) / np.linalg.norm(adref)
        assert eb < 2e-6 and ea < 1e-6, (r, eb, ea)
    assert not np.any(ao)
