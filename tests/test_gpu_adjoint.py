"""Device adjoint of Q_abs (ddscat_hip_cabs_adjoint) and the torch autograd
wrapper against the FP64 dense oracle adjoint (itself pinned by central
differences in test_oracle_adjoint.py) and against a directional central
difference of the device forward model.

Bars: device dCabs/d(adia) within 2e-3 relative L2 of the FP64 oracle
adjoint on the device solution (solver tolerance 1e-6); autograd dJ/d eps
within 2e-3 of the oracle chain rule; directional central difference (step
1e-3 on eps, as the inverse-design check) within 1e-2 of the autograd
directional derivative."""
import math

import numpy as np
import pytest

from oracle import ddscat_oracle as O
from oracle import adjoint as ADJ
from ddscat_hip import targets as T

pytestmark = pytest.mark.gpu

TOL = 1e-6


@pytest.fixture(scope="module")
def design():
    tg = T.sphere(8)
    kd = T.kd_for_size_parameter(tg, 1.5)
    rng = np.random.default_rng(44)
    eps = rng.uniform(1.0, 4.0, tg.nat0) + 0.05j
    khat = np.array([0.0, 0.0, 1.0])
    pols = list(T.polarization_pair(khat))
    return dict(tg=tg, kd=kd, eps=eps, khat=khat, pols=pols)


def prep(lib, d):
    tg = d["tg"]
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, d["khat"] * d["kd"])


def oracle_gradients(d):
    """Per polarization: (g_adia [3 nat0], P, Ao) from the FP64 dense model."""
    tg, kd = d["tg"], d["kd"]
    kd32 = float(np.float32(kd))
    out = []
    for e0 in d["pols"]:
        ai = 1.0 / T.ldr_alpha(d["eps"], kd, d["khat"], e0)
        ad = np.tile(ai.astype(np.complex64).astype(complex), 3)
        ao = np.zeros_like(ad)
        Ao = ADJ.compact_dense_operator(O.dense_operator, tg.nx, tg.ny, tg.nz, kd32, ad, ao, tg.iocc)
        ef = T.incident_field(tg, d["khat"] * kd, e0, tg.centroid_x0())
        e = ef.reshape(3, -1)[:, tg.occ].reshape(-1).astype(complex)
        _, p = ADJ.cabs_dense(Ao, e, ad, ao, kd32)
        ga, go = ADJ.cabs_adjoint(Ao, p, ad, ao, kd32)
        out.append((ga, go, p, ad, e))
    return out


@pytest.mark.parametrize("backend", ["fft3d", "slices"])
def test_device_adjoint_matches_oracle(lib, design, backend):
    """dCabs/d(adia, aoff) from the forward and adjoint batched solves on
    either MATVEC backend against the FP64 dense adjoint."""
    from ddscat_hip import BACKEND_FFT3D, BACKEND_SLICES
    lib.release()
    lib.set_backend(BACKEND_SLICES if backend == "slices" else BACKEND_FFT3D)
    try:
        _device_adjoint_case(lib, design)
    finally:
        lib.release()
        lib.set_backend(BACKEND_FFT3D)


def _device_adjoint_case(lib, design):
    import torch
    d = design
    prep(lib, d)
    tg, kd = d["tg"], d["kd"]
    ora = oracle_gradients(d)
    B = len(ora)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(np.stack(a)).astype(np.complex64)).cuda()
    ad = dev([o[3] for o in ora])
    ao = torch.zeros_like(ad)
    e = dev([o[4] for o in ora])
    x = torch.zeros_like(e)
    torch.cuda.synchronize()
    lib.solve_batch("qmrccg", B, e.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(), maxit=500, tol=TOL)
    ga, go = torch.empty_like(ad), torch.empty_like(ad)
    ak = np.tile((d["khat"] * kd).astype(np.float32), (B, 1))
    it, te, rc = lib.cabs_adjoint("qmrccg", B, x.data_ptr(), ad.data_ptr(), ao.data_ptr(), ak, ga.data_ptr(),
                                  go.data_ptr(), maxit=500, tol=TOL)
    ga, go = ga.cpu().numpy(), go.cpu().numpy()
    for r, (g_ref, go_ref, p_ref, _, _) in enumerate(ora):
        ea = np.linalg.norm(ga[r] - g_ref) / np.linalg.norm(g_ref)
        eo = np.linalg.norm(go[r] - go_ref) / np.linalg.norm(go_ref)
        print(f"pol {r}: adjoint it={it[r]} tole={te[r]:.2e}  rel err g_adia={ea:.2e} g_aoff={eo:.2e}")
        assert ea <= 2e-3 and eo <= 2e-3


def test_autograd_eps_gradient_matches_oracle_chain(lib, design):
    import torch
    from ddscat_hip.adjoint import design_q_abs
    d = design
    prep(lib, d)
    eps = torch.tensor(d["eps"], dtype=torch.complex128, requires_grad=True)
    J = design_q_abs(lib, d["tg"], eps, d["kd"], d["khat"], d["pols"], tol=TOL)
    J.backward()
    got = eps.grad.numpy()
    # oracle: dJ = Re(sum_c g_c a'(eps) d eps) / (geo * npol)  ->  torch grad = conj(...)
    tg, kd = d["tg"], d["kd"]
    geo = math.pi * (3.0 * tg.nat0 / (4.0 * math.pi)) ** (2.0 / 3.0)
    ref = np.zeros(tg.nat0, complex)
    ora = oracle_gradients(d)
    for (ga, _, _, _, _), e0 in zip(ora, d["pols"]):
        khat = d["khat"]
        ea = np.abs(e0)
        S = float(np.sum((khat * ea / np.linalg.norm(ea)) ** 2))
        dprime = -4 * np.pi / (d["eps"] - 1) ** 2 + (T.LDR_B2 + T.LDR_B3 * S) * kd ** 2
        ref += np.conj(ga.reshape(3, -1).sum(axis=0) * dprime)
    ref /= geo * len(ora)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"dJ/deps rel err vs FP64 oracle chain: {err:.2e}; J = {J.item():.6f}")
    assert err <= 2e-3


def test_autograd_directional_central_difference(lib, design):
    import torch
    from ddscat_hip.adjoint import design_q_abs
    d = design
    prep(lib, d)
    rng = np.random.default_rng(9)
    v = rng.normal(size=d["tg"].nat0) + 1j * rng.normal(size=d["tg"].nat0)
    v /= np.abs(v).max()
    eps = torch.tensor(d["eps"], dtype=torch.complex128, requires_grad=True)
    J = design_q_abs(lib, d["tg"], eps, d["kd"], d["khat"], d["pols"], tol=TOL)
    J.backward()
    ana = float(np.sum(np.real(np.conj(eps.grad.numpy()) * v)))
    h = 1e-3
    with torch.no_grad():
        jp = design_q_abs(lib, d["tg"], torch.tensor(d["eps"] + h * v), d["kd"], d["khat"], d["pols"], tol=TOL)
        jm = design_q_abs(lib, d["tg"], torch.tensor(d["eps"] - h * v), d["kd"], d["khat"], d["pols"], tol=TOL)
    num = (jp.item() - jm.item()) / (2 * h)
    print(f"directional derivative: adjoint {ana:.6e}  central difference {num:.6e}")
    assert abs(num - ana) <= 1e-2 * abs(ana)
