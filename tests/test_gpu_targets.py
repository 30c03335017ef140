"""Several prepared targets at once (target slots) and their batched solves in
one call (ddscat_hip_solve_targets: one stream and host thread per target).

The first three targets of the seeded uneven batch (72^3, 72^3, 96^3 grids),
2 orientations x 2 polarizations each, are prepared in slots 1-3 and their
inputs built on the device.  The concurrent multi-target PBCGS2 must give the
same solutions, bit for bit, as batched solves of one target after another,
and Q_ext / Q_abs within 1e-4 of the FP64 oracle (tests/golden/
config_ellipsoids.npz).  Slot 0, the reference ABI's target, keeps its own
state meanwhile.
"""
from pathlib import Path

import numpy as np
import pytest

from ddscat_hip import PREC_MIXED
import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden"
TOL = 1e-5


def test_concurrent_targets_match_sequential_and_oracle(lib):
    import torch
    f = np.load(GOLDEN / "config_ellipsoids.npz", allow_pickle=False)
    cases = W.ellipsoid_cases()
    lib.set_precision(PREC_MIXED)
    # slot 0 holds a small target of its own through the test
    from ddscat_hip import targets as T
    tg0 = T.sphere(8)
    lib.select_target(0)
    lib.prepare_green(tg0.nx, tg0.ny, tg0.nz, tg0.iocc, (0.2, 0.0, 0.0))
    info0 = lib.info()
    bufs = []
    try:
        for t, case in enumerate(cases):
            tg = case["tg"]
            lib.select_target(t + 1)
            lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, case["aks"][0])
            B, n = len(case["aks"]), 3 * tg.nat0
            eps = torch.full((tg.nat0,), complex(case["m"]) ** 2, dtype=torch.complex64, device="cuda")
            b = torch.empty((B, n), dtype=torch.complex64, device="cuda")
            ad, ao = torch.empty_like(b), torch.empty_like(b)
            lib.setup_rhs(case["aks"], case["e0s"], tg.centroid_x0(), eps.data_ptr(), b.data_ptr(), ad.data_ptr(),
                          ao.data_ptr())
            bufs.append(dict(b=b, ad=ad, ao=ao, B=B))
        # one target after another
        seq = []
        for t, d in enumerate(bufs):
            lib.select_target(t + 1)
            x = torch.zeros_like(d["b"])
            lib.solve_batch("pbcgs2", d["B"], d["b"].data_ptr(), x.data_ptr(), d["ad"].data_ptr(),
                            d["ao"].data_ptr(), maxit=500, tol=TOL)
            seq.append(x)
        # all targets in one call
        xs = [torch.zeros_like(d["b"]) for d in bufs]
        per, rc = lib.solve_targets("pbcgs2", [1, 2, 3], [d["B"] for d in bufs], [d["b"].data_ptr() for d in bufs],
                                    [x.data_ptr() for x in xs], [d["ad"].data_ptr() for d in bufs],
                                    [d["ao"].data_ptr() for d in bufs], maxit=500, tol=TOL)
        torch.cuda.synchronize()
        assert rc == 0
        for t, (case, d) in enumerate(zip(cases, bufs)):
            it, te, nc = per[t]
            assert np.all(te <= TOL), te
            assert torch.equal(xs[t], seq[t]), f"target {t}: concurrent solve differs from the sequential one"
            lib.select_target(t + 1)
            q = lib.evalq_batch(d["B"], d["b"].data_ptr(), xs[t].data_ptr(), d["ad"].data_ptr(), d["ao"].data_ptr(),
                                case["aks"])
            geo = W.geometric_cross_section(case["tg"])
            e_ext = np.max(np.abs(q[:, 1] / geo - f[f"t{t}_qext"]) / np.abs(f[f"t{t}_qext"]))
            e_abs = np.max(np.abs(q[:, 0] / geo - f[f"t{t}_qabs"]) / np.abs(f[f"t{t}_qabs"]))
            print(f"slot {t + 1} ({case['tg'].nx}^3-ish, {case['tg'].nat0} dipoles): iterations {it.tolist()}, "
                  f"Qext rel err {e_ext:.2e}, Qabs rel err {e_abs:.2e}")
            assert e_ext <= 1e-4 and e_abs <= 1e-4
        lib.select_target(0)
        assert lib.info()["nat"] == info0["nat"] == tg0.nat
    finally:
        for t in range(len(cases)):
            lib.select_target(t + 1)
            lib.release()
        lib.select_target(0)
        lib.release()
