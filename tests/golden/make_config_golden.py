"""Writes the FP64-oracle fixtures of the full-size configuration tests.

    python tests/golden/make_config_golden.py [block128] [ellipsoids] [design64] [sphere64]

Each workload (tests/workloads.py) is solved here, in complex128, by the
oracle's pruned FFT MATVEC (oracle.ddscat_oracle.FftMatvec64, equal to the
dense closed-form operator on small grids) and the oracle's QMRCCG
(oracle.solvers.qmrccg) at a residual of 1e-8, three orders below the 1e-5
the device solves run at.  The fixtures hold only small results (cross
sections, iteration counts, the gradient map of the 64^3 design) and a
checksum of the regenerated inputs.  The reference repository ships no
expected outputs for these workloads; these numbers are the oracle's.

Run time on 8 host cores: about 20 min (block128), 6 min (ellipsoids),
6 min (design64), 10 min (sphere64).
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
for p in (ROOT, ROOT / "ddscat-cuda_amd", HERE.parent):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

from oracle import ddscat_oracle as O  # noqa: E402
from oracle import solvers as S  # noqa: E402
from oracle import adjoint as OA  # noqa: E402
import workloads as W  # noqa: E402

TOL = 1e-8
WORKERS = 8


def log(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)


def solve(op, b, adia, aoff, maxit=3000):
    mv = lambda cw, v: op(cw, v, adia, aoff)
    x, it, res, nc = S.qmrccg(mv, b, maxit=maxit, tol=TOL)
    true_res = np.linalg.norm(b - op("N", x, adia, aoff)) / np.linalg.norm(b)
    assert true_res < 3 * TOL, true_res
    return x, it, true_res


def block128():
    single_target(W.block128(), "config_block128.npz")


def sphere64():
    single_target(W.sphere64(), "config_sphere64.npz")


def single_target(w, fname):
    tg, kd, khat = w["tg"], w["kd"], w["khat"]
    ak = (khat * kd).astype(np.float32)
    akd32 = float(np.linalg.norm(ak.astype(np.float64)))
    log(fname, "green octant", tg.nx)
    go = O.green_octant_separable(tg.nx, tg.ny, tg.nz, akd32)
    op = O.FftMatvec64(go, tg.nx, tg.ny, tg.nz, tg.iocc, workers=WORKERS)
    geo = W.geometric_cross_section(tg)
    out = dict(qext=[], qabs=[], qpha=[], iterations=[], true_residual=[])
    for k, (b, adia, aoff) in enumerate(w["rhs"]):
        log(fname, "solve polarization", k)
        x, it, tr = solve(op, b.astype(np.complex128), adia.astype(np.complex128), aoff.astype(np.complex128))
        cabs, cext, cpha = O.evalq(b, x, adia, aoff, ak, 1.0)
        out["qext"].append(cext / geo); out["qabs"].append(cabs / geo); out["qpha"].append(cpha / geo)
        out["iterations"].append(it); out["true_residual"].append(tr)
        log(f"{fname}: pol {k}: {it} iterations, residual {tr:.2e}, Qext {cext / geo:.9f} Qabs {cabs / geo:.9f}")
    sums = W.checksum(*[a for r in w["rhs"] for a in r])
    np.savez(HERE / fname, checksum=np.array(sums), tol=np.array(TOL), **{k: np.array(v) for k, v in out.items()})
    log("wrote", fname)


def ellipsoids():
    cases = W.ellipsoid_cases()
    res = {}
    for ci, case in enumerate(cases):
        tg = case["tg"]
        akd32 = float(np.linalg.norm(case["aks"][0].astype(np.float64)))
        log(f"ellipsoids: target {ci} grid {tg.nx}x{tg.ny}x{tg.nz} nat0 {tg.nat0} m {case['m']:.4f}")
        go = O.green_octant_separable(tg.nx, tg.ny, tg.nz, akd32)
        op = O.FftMatvec64(go, tg.nx, tg.ny, tg.nz, tg.iocc, workers=WORKERS)
        geo = W.geometric_cross_section(tg)
        qext, qabs, its = [], [], []
        for r, (b, adia, aoff) in enumerate(W.ellipsoid_rhs_host(case)):
            x, it, tr = solve(op, b.astype(np.complex128), adia, aoff)
            cabs, cext, _ = O.evalq(b, x, adia, aoff, case["aks"][r], 1.0)
            qext.append(cext / geo); qabs.append(cabs / geo); its.append(it)
            log(f"ellipsoids: target {ci} rhs {r}: {it} iterations, Qext {cext / geo:.8f} Qabs {cabs / geo:.8f}")
        res[f"t{ci}_qext"] = np.array(qext)
        res[f"t{ci}_qabs"] = np.array(qabs)
        res[f"t{ci}_iterations"] = np.array(its)
        res[f"t{ci}_checksum"] = np.array(W.checksum(tg.iocc, case["aks"], case["e0s"]))
    np.savez(HERE / "config_ellipsoids.npz", tol=np.array(TOL), ntargets=np.array(len(cases)), **res)
    log("wrote config_ellipsoids.npz")


def design64():
    w = W.design64()
    tg, kd, khat, eps, geo = w["tg"], w["kd"], w["khat"], w["eps"], w["geo"]
    ak = (khat * kd).astype(np.float32)
    akk = float(np.linalg.norm(ak.astype(np.float64)))
    go = O.green_octant_separable(tg.nx, tg.ny, tg.nz, akk)
    op = O.FftMatvec64(go, tg.nx, tg.ny, tg.nz, tg.iocc, workers=WORKERS)
    qabs, grad, its = [], np.zeros(tg.nat0, np.complex128), []
    for k, e0 in enumerate(w["pols"]):
        inv, dinv = W.ldr_inverse(eps, kd, khat, e0)
        adia = np.tile(inv, 3)
        aoff = np.zeros_like(adia)
        b = W.incident_compact(tg, khat * kd, e0).astype(np.complex128)
        log(f"design64: forward solve, polarization {k}")
        p, it, _ = solve(op, b, adia, aoff)
        cabs = OA.cabs_of(p, adia, aoff, akk)
        log(f"design64: adjoint solve, polarization {k} (forward {it} iterations, Qabs {cabs / geo:.9f})")
        adj_solve = lambda rhs: solve(op, rhs, adia, aoff)[0]
        g_adia, _ = OA.cabs_adjoint(adj_solve, p, adia, aoff, akk)
        # J = mean over polarizations of Cabs / geo; every component's 1/alpha is inv(eps)
        grad += g_adia.reshape(3, -1).sum(axis=0) * dinv / (geo * len(w["pols"]))
        qabs.append(cabs / geo); its.append(it)
    np.savez_compressed(HERE / "config_design64.npz", tol=np.array(TOL), qabs=np.array(qabs),
                        iterations=np.array(its), grad_eps=grad.astype(np.complex64),
                        checksum=np.array(W.checksum(eps)))
    log(f"wrote config_design64.npz (J = {np.mean(qabs):.9f})")


if __name__ == "__main__":
    todo = sys.argv[1:] or ["ellipsoids", "design64", "block128"]
    for name in todo:
        globals()[name]()
