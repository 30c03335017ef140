"""GPU Krylov solvers, EVALQ and SCAT against the oracle.

Bars: solution relative error vs the dense direct solve <= 30 tol (the
solvers stop on the relative residual, cond(A) ~ 10 here); true residual
(FP64, dense) <= 2 tol; Q_ext/Q_abs within 1e-4 relative of the FP64 oracle
at tol 1e-5; EVALQ/SCAT on the same polarization within 1e-6 relative."""
import math

import numpy as np
import pytest

from oracle import ddscat_oracle as O
from oracle import solvers as S
from ddscat_hip import targets as T

pytestmark = pytest.mark.gpu

ALL = ["pbcgs2", "qmrccg", "petrkp", "gpbicg", "pbcgst", "bicgstab4", "gpbicgstab2", "gpbicgstab4"]


def problem(D, m, x=1.0, khat=(1.0, 0.0, 0.0), pol=(0, 1, 0)):
    tg = T.sphere(D)
    kd = T.kd_for_size_parameter(tg, x)
    khat = np.asarray(khat, float)
    e0 = np.asarray(pol, complex)
    adia, aoff = T.operators(tg, m, kd, khat, e0)
    x0 = tg.centroid_x0()
    b = T.incident_field(tg, khat * kd, e0, x0)
    return tg, kd, khat, e0, adia, aoff, b, x0


def dense(tg, kd, adia, aoff):
    kd32 = float(np.float32(kd))
    A = O.dense_operator(tg.nx, tg.ny, tg.nz, kd32, adia.astype(complex), aoff.astype(complex), tg.iocc)
    occ3 = np.flatnonzero(np.tile(tg.iocc != 0, 3))
    return A[np.ix_(occ3, occ3)], occ3


@pytest.fixture(scope="module")
def small(lib):
    tg, kd, khat, e0, adia, aoff, b, x0 = problem(8, 1.33 + 0.01j)
    Ao, occ3 = dense(tg, kd, adia, aoff)
    xref = np.linalg.solve(Ao, b.astype(complex)[occ3])
    return dict(tg=tg, kd=kd, khat=khat, e0=e0, adia=adia, aoff=aoff, b=b, x0=x0, Ao=Ao, occ3=occ3, xref=xref)


@pytest.mark.parametrize("name", ALL)
def test_solver_matches_dense_solution(lib, small, name):
    P = small
    tg = P["tg"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, P["khat"] * P["kd"])
    tol = 1e-5
    x, it, tole, nc, rc = lib.solve(name, P["b"], P["adia"], P["aoff"], maxit=300, tol=tol)
    xo = x.astype(complex)[P["occ3"]]
    bo = P["b"].astype(complex)[P["occ3"]]
    true_res = np.linalg.norm(P["Ao"] @ xo - bo) / np.linalg.norm(bo)
    err = np.linalg.norm(xo - P["xref"]) / np.linalg.norm(P["xref"])
    print(f"{name}: it={it} matvecs={nc} tole={tole:.2e} true={true_res:.2e} err={err:.2e}")
    assert rc == 0
    assert true_res <= 2 * tol
    assert err <= 30 * tol
    assert tole <= 2 * tol
    assert np.all(x.reshape(3, -1)[:, tg.iocc == 0] == 0)


@pytest.mark.parametrize("name", ALL)
def test_solver_ignores_uninitialised_workspace(lib, small, name):
    """Work vectors that held other data (here NaN bit patterns in the whole
    workspace) must not leak into a solve: every solver writes a vector,
    zeros included, before it reads it.  Regression: clearing a vector as
    0 * itself kept NaN and broke BiCGStab(4)'s MR polynomial."""
    P = small
    tg = P["tg"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, P["khat"] * P["kd"])
    lib.poison_workspace()
    tol = 1e-5
    x, it, tole, nc, rc = lib.solve(name, P["b"], P["adia"], P["aoff"], maxit=300, tol=tol)
    xo = x.astype(complex)[P["occ3"]]
    bo = P["b"].astype(complex)[P["occ3"]]
    true_res = np.linalg.norm(P["Ao"] @ xo - bo) / np.linalg.norm(bo)
    print(f"{name} on a NaN-filled workspace: it={it} tole={tole:.2e} true={true_res:.2e}")
    assert rc == 0 and np.all(np.isfinite(x))
    assert true_res <= 2 * tol


@pytest.mark.parametrize("name", ["pbcgs2", "qmrccg", "petrkp"])
def test_iteration_count_band_vs_oracle(lib, small, name):
    """Same algorithm: device and FP64 oracle iteration counts agree loosely."""
    P = small
    tg = P["tg"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, P["khat"] * P["kd"])
    _, it, _, _, _ = lib.solve(name, P["b"], P["adia"], P["aoff"], maxit=300, tol=1e-5)
    Ao = P["Ao"]
    mv = lambda cw, v: Ao @ v if cw == "N" else np.conj(Ao @ np.conj(v))
    _, it_ref, _, _ = getattr(S, name)(mv, P["b"].astype(complex)[P["occ3"]], maxit=300, tol=1e-5)
    print(f"{name}: device it={it}, oracle it={it_ref}")
    assert abs(it - it_ref) <= max(2, 0.25 * it_ref)


def test_evalq_scat_match_oracle_on_same_polarization(lib, small):
    P = small
    tg = P["tg"]
    ak = P["khat"] * P["kd"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, ak)
    x, *_ = lib.solve("pbcgs2", P["b"], P["adia"], P["aoff"], maxit=300, tol=1e-6)
    cabs, cext, cpha, rc = lib.evalq(ak, 1.0)
    occ = tg.occ
    comp = lambda v: v.reshape(3, -1)[:, occ].reshape(-1)
    ref = O.evalq(comp(P["b"]), comp(x), comp(P["adia"]), comp(P["aoff"]), np.float32(ak), 1.0)
    for got, want in zip((cabs, cext, cpha), ref):
        assert abs(got - want) <= 1e-6 * abs(want) + 1e-12
    # non-resident path with the same compact arrays
    c2 = lib.evalq(ak, 1.0, use_resident=False, e_c=comp(P["b"]), p_c=comp(x), adia_c=comp(P["adia"]),
                   aoff_c=comp(P["aoff"]))
    assert np.allclose(c2[:3], (cabs, cext, cpha), rtol=1e-12)
    # SCAT: integration + two selected directions
    aks = np.array([[0, 0, 1], [0, 1, 0]], np.float32) * np.float32(P["kd"])
    em1 = np.array([[1, 0, 0], [1, 0, 0]], np.float32)
    em2 = np.array([[0, 1, 0], [0, 0, 1]], np.float32)
    x0 = np.float32(P["x0"])
    e01 = P["e0"]
    res = lib.scat(ak, (1, 1, 1), x0, e01, 0.5, aks=aks, em1=em1, em2=em2)
    want = O.scat(comp(x), occ, tg.nx, tg.ny, np.float32(ak), (1.0, 1.0, 1.0), x0.astype(float), e01, 0.5, 1.0,
                  aks=aks, em1=em1, em2=em2)
    assert res["navg"] == want["navg"]
    for k in ("csca", "cscag2", "cbksca"):
        assert abs(res[k] - want[k]) <= 1e-6 * abs(want[k]) + 1e-14, k
    assert np.allclose(res["cscag"], want["cscag"], rtol=1e-6, atol=1e-9 * abs(want["csca"]))
    assert np.allclose(res["f1"], want["f1"], rtol=1e-5, atol=1e-7 * np.abs(want["f1"]).max())
    assert np.allclose(res["f2"], want["f2"], rtol=1e-5, atol=1e-7 * np.abs(want["f2"]).max())
    # optical theorem on a converged solution: Cext = Cabs + Csca (quadrature error only)
    assert abs(cext - (cabs + res["csca"])) < 0.02 * cext


def test_sphere_d32_pbcgs2_qext_vs_fp64_oracle(lib):
    """Small workload: sphere D=32 (17,256 dipoles), m=1.33+0.01i, x=2, PBCGS2
    tol 1e-5, both polarizations: Q_ext / Q_abs within 1e-4 of the FP64 oracle solve."""
    tg, kd, khat, _, _, _, _, x0 = problem(32, 1.33 + 0.01j, x=2.0)
    ak = khat * kd
    kd32 = float(np.float32(kd))
    goct = O.green_octant(tg.nx, tg.ny, tg.nz, kd32)
    occ = tg.occ
    pia2 = math.pi * tg.a_eff() ** 2
    for pol in ((0, 1, 0), (0, 0, 1)):
        e0 = np.asarray(pol, complex)
        adia, aoff = T.operators(tg, 1.33 + 0.01j, kd, khat, e0)
        b = T.incident_field(tg, ak, e0, x0)
        lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, ak)
        x, it, tole, nc, rc = lib.solve("pbcgs2", b, adia, aoff, maxit=300, tol=1e-5)
        cabs, cext, cpha, _ = lib.evalq(ak, 1.0)
        comp = lambda v: np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)
        ad_c, ao_c, b_c = comp(adia).astype(complex), comp(aoff).astype(complex), comp(b).astype(complex)

        def mv(cw, v):
            full = O.scatter_compact(v, occ, tg.nat)
            return comp(O.matvec_fft(cw, full, adia.astype(complex), aoff.astype(complex), goct, tg.nx, tg.ny,
                                     tg.nz, tg.iocc))
        xr, itr, resr, _ = S.pbcgs2(mv, b_c, maxit=300, tol=1e-8)
        ref = O.evalq(b_c, xr, ad_c, ao_c, np.float32(ak), 1.0)
        qabs, qext = cabs / pia2, cext / pia2
        print(f"pol {pol}: it={it} tole={tole:.2e} Qext={qext:.6f} (oracle {ref[1]/pia2:.6f}) "
              f"Qabs={qabs:.6f} (oracle {ref[0]/pia2:.6f})")
        assert abs(cext - ref[1]) <= 1e-4 * abs(ref[1])
        assert abs(cabs - ref[0]) <= 1e-4 * abs(ref[0])


def test_solver_error_codes(lib, small):
    P = small
    tg = P["tg"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, P["khat"] * P["kd"])
    n = 3 * tg.nat
    b = P["b"]
    z = np.zeros(n, np.complex64)
    it, te, nc = (np.zeros(1, np.int32), np.zeros(1, np.float32), np.zeros(1, np.int32))
    p = lambda a: a.ctypes.data
    codes = {"gpbicg": 30, "qmrccg": 50, "pbcgst": 70, "pbcgs2": 90, "petrkp": 110, "bicgstab4": 130,
             "gpbicgstab2": 150, "gpbicgstab4": 180}
    for name, code in codes.items():
        fn = getattr(lib.lib, f"ddscat_cuda_solve_{name}_f32")
        assert fn(p(b), p(z), p(P["adia"]), p(P["aoff"]), n - 3, 10, 1e-5, p(it), p(te), p(nc)) == code
        assert fn(p(b), p(z), p(P["adia"]), p(P["aoff"]), n, 10, 0.0, p(it), p(te), p(nc)) == code
    # zero right-hand side
    assert lib.solve("qmrccg", z, P["adia"], P["aoff"], check=False)[4] == 52
    # MAXIT: PBCGST reports 87, PBCGS2 returns 0 with the achieved residual
    assert lib.solve("pbcgst", b, P["adia"], P["aoff"], maxit=2, tol=1e-12, check=False)[4] == 87
    x, it2, tole, nc2, rc = lib.solve("pbcgs2", b, P["adia"], P["aoff"], maxit=2, tol=1e-12, check=False)
    assert rc == 0 and it2 == 2 and tole > 1e-12
    # EVALQ / SCAT without a resident solution
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, P["khat"] * P["kd"])
    assert lib.evalq(P["khat"] * P["kd"], 1.0, check=False)[3] == 301
    assert lib.scat(P["khat"] * P["kd"], (1, 1, 1), (0, 0, 0), P["e0"], 0.5, check=False)["rc"] == 320


@pytest.mark.parametrize("name", ["pbcgs2", "qmrccg", "gpbicgstab2"])
def test_solve_device_matches_host_entry(lib, small, name):
    """ddscat_hip_solve_device on compact device vectors runs the same solver
    as the reference entry point: same iterate, iteration and MATVEC counts."""
    import torch
    P = small
    tg = P["tg"]
    occ = tg.occ
    comp = lambda v: np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)).astype(np.complex64)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, P["khat"] * P["kd"])
    xh, it_h, tole_h, nc_h, _ = lib.solve(name, P["b"], P["adia"], P["aoff"], maxit=300, tol=1e-5)
    b, ad, ao = dev(comp(P["b"])), dev(comp(P["adia"])), dev(comp(P["aoff"]))
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    it_d, tole_d, nc_d, rc = lib.solve_device(name, b.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                              maxit=300, tol=1e-5)
    xd = x.cpu().numpy().view(np.complex64)
    print(f"{name}: host it={it_h} nc={nc_h}  device it={it_d} nc={nc_d}")
    assert rc == 0 and it_d == it_h and nc_d == nc_h
    assert np.array_equal(xd, comp(xh))
    assert abs(tole_d - tole_h) <= 1e-12 + 1e-6 * tole_h
    cabs, cext, _, _ = lib.evalq(P["khat"] * P["kd"])
    assert cext > 0 and cabs > 0


@pytest.mark.parametrize("name,mode", [("pbcgs2", "restart"), ("qmrccg", "restart"), ("qmrccg", "replace"),
                                       ("gpbicg", "restart"), ("bicgstab4", "restart")])
def test_forced_reliable_restart_still_converges(lib, small, name, mode):
    """The reliable-residual policy (cu:543-556): a gap >= 1e-3 at a check
    restarts the recursion from the true residual (QMRCCG in the reference's
    restart mode; in its default mode it replaces the recursive residual and
    keeps the recurrence).  The hook forces one such gap; the solve must act
    on it and still reach the dense solution."""
    P = small
    tg = P["tg"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, P["khat"] * P["kd"])
    tol = 1e-6
    lib.set_reliable_mode(mode)
    try:
        x0, it0, *_ = lib.solve(name, P["b"], P["adia"], P["aoff"], maxit=300, tol=tol)
        lib.force_restart()
        x, it, tole, nc, rc = lib.solve(name, P["b"], P["adia"], P["aoff"], maxit=300, tol=tol)
    finally:
        lib.set_reliable_mode("replace")
    restarts = lib.info()["restarts"]
    actions = lib.solve_log(0)["checks"][:, 3].astype(int)
    xo = x.astype(complex)[P["occ3"]]
    err = np.linalg.norm(xo - P["xref"]) / np.linalg.norm(P["xref"])
    print(f"{name} ({mode}): restarts={restarts} actions={actions.tolist()} it={it} (unforced {it0}) err={err:.2e}")
    assert rc == 0
    if mode == "restart":
        assert restarts >= 1 and np.any(actions == 1)
    else:
        assert restarts == 0 and np.any(actions == 2)
    assert err <= 30 * tol


def test_iocc_count_mismatch_codes(lib):
    """IOCC with a different number of occupied sites than NAT0: prepare_green
    returns 208 and prepare returns 6 (cu:688, cu:1016, cu:1025)."""
    nx, ny, nz = 4, 4, 4
    iocc = np.ones(nx * ny * nz, np.int16)
    iocc[:5] = 0
    assert lib.prepare_green(nx, ny, nz, iocc, (0.2, 0, 0), nat0=int(iocc.sum()) + 1, check=False) == 208
    g = np.zeros(6 * (nx + 1) * (ny + 1) * (nz + 1), np.complex64)
    assert lib.prepare(nx, ny, nz, g, iocc, nat0=int(iocc.sum()) - 1, check=False) == 6
    lib.release()
