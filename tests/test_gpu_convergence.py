"""QMRCCG on the large-target material: sphere D=64, m = 1.7+0.05i, kd = 0.25
(the D=384 workload's |m| kd = 0.43 at 1/216 of its dipoles).

With FP32 vectors the recursive and true residuals drift apart by about
1e-6 |b|.  Below a residual of about 1e-3 every reliable check
(cu:543-556, every 20 iterations) then sees a gap >= 1e-3.  The reference's
restart of the Lanczos recurrence at each such check stalls the solve, so
by default the library replaces the recursive residual with the true one
instead (ddscat_hip_set_reliable_mode).  Checked here:
  * batched mixed QMRCCG converges to 1e-5 with no restart, within 2x the
    iterations of the FP64 oracle QMRCCG on the same inputs, and Q_abs /
    Q_ext of both polarizations (equal by the lattice's symmetry) match the
    oracle's 1e-8 solution within 1e-4;
  * the single-RHS solver through the device entry does the same;
  * the reference's restart policy, still available, restarts at most
    checks and needs many times the iterations (the behaviour that stalled
    the D=384 solve).
"""
from pathlib import Path

import numpy as np
import pytest

from ddscat_hip import PREC_MIXED
import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden"
TOL = 1e-5


@pytest.fixture(scope="module")
def sphere(lib):
    import torch
    w = W.sphere64()
    tg = w["tg"]
    lib.set_precision(PREC_MIXED)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, w["khat"] * w["kd"])
    w["b"] = torch.from_numpy(np.stack([r[0] for r in w["rhs"]])).cuda()
    w["ad"] = torch.from_numpy(np.stack([r[1] for r in w["rhs"]])).cuda()
    w["ao"] = torch.from_numpy(np.stack([r[2] for r in w["rhs"]])).cuda()
    w["ak"] = np.tile((w["khat"] * w["kd"]).astype(np.float32), (2, 1))
    w["fix"] = np.load(GOLDEN / "config_sphere64.npz", allow_pickle=False)
    assert str(w["fix"]["checksum"]) == W.checksum(*[a for r in w["rhs"] for a in r]), "inputs drifted"
    yield w
    lib.set_reliable_mode("replace")
    lib.release()


def batched(lib, w, mode, maxit):
    import torch
    lib.set_reliable_mode(mode)
    x = torch.zeros_like(w["b"])
    it, te, nc, rc = lib.solve_batch("qmrccg", 2, w["b"].data_ptr(), x.data_ptr(), w["ad"].data_ptr(),
                                     w["ao"].data_ptr(), maxit=maxit, tol=TOL)
    q = lib.evalq_batch(2, w["b"].data_ptr(), x.data_ptr(), w["ad"].data_ptr(), w["ao"].data_ptr(), w["ak"])
    logs = [lib.solve_log(r) for r in range(2)]
    return it, te, q, logs


def test_qmrccg_converges_with_residual_replacement(lib, sphere):
    w, f = sphere, sphere["fix"]
    it, te, q, logs = batched(lib, w, "replace", 1000)
    geo = W.geometric_cross_section(w["tg"])
    acts = [l["checks"][:, 3].astype(int) for l in logs]
    print(f"replace: iterations {it.tolist()} (FP64 oracle to 1e-8: {f['iterations'].tolist()}), residual "
          f"{te.tolist()}, restarts {[l['restarts'] for l in logs]}, replacements {[int(np.sum(a == 2)) for a in acts]}")
    assert np.all(te <= TOL)
    assert all(l["restarts"] == 0 for l in logs)
    assert np.all(it <= 2 * np.max(f["iterations"]))
    for k in range(2):
        assert abs(q[k, 0] / geo - f["qabs"][k]) <= 1e-4 * f["qabs"][k]
        assert abs(q[k, 1] / geo - f["qext"][k]) <= 1e-4 * f["qext"][k]
    assert abs(q[0, 0] - q[1, 0]) <= 1e-4 * q[0, 0]


def test_single_rhs_qmrccg_converges(lib, sphere):
    import torch
    w = sphere
    lib.set_reliable_mode("replace")
    x = torch.zeros_like(w["b"][0])
    it, te, nc, rc = lib.solve_device("qmrccg", w["b"][0].data_ptr(), x.data_ptr(), w["ad"][0].data_ptr(),
                                      w["ao"][0].data_ptr(), maxit=1000, tol=TOL)
    lg = lib.solve_log(0)
    print(f"single-RHS QMRCCG: {it} iterations, residual {te:.2e}, restarts {lg['restarts']}")
    assert te <= TOL and lg["restarts"] == 0 and it <= 2 * np.max(w["fix"]["iterations"])


def test_reference_restart_policy_stalls(lib, sphere):
    it_r, te_r, _, logs = batched(lib, sphere, "restart", 600)
    lib.set_reliable_mode("replace")
    restarts = [l["restarts"] for l in logs]
    checks = [len(l["checks"]) for l in logs]
    print(f"restart policy: iterations {it_r.tolist()}, residual {te_r.tolist()}, restarts {restarts} of "
          f"{checks} checks")
    # at 600 iterations the restarted recurrence is far from done; most checks restarted
    assert np.any(te_r > TOL) or np.all(it_r > 500)
    assert max(restarts) >= 10
