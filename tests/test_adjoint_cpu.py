"""Host-side pieces of the differentiable path (no GPU): the torch LDR map
equals 1/targets.ldr_alpha and its autograd derivative equals the closed
form d(1/alpha)/d eps = -4 pi/(eps-1)^2 + (b2 + b3 S) kd^2."""
import math

import numpy as np
import torch

from ddscat_hip import targets as T
from ddscat_hip.adjoint import ldr_alpha_inv


def test_ldr_alpha_inv_matches_numpy_ldr():
    rng = np.random.default_rng(0)
    eps = rng.uniform(1, 4, 50) + 1j * rng.uniform(0, 0.3, 50)
    khat = np.array([0.3, 0.4, 0.866]); khat /= np.linalg.norm(khat)
    for ehat in T.polarization_pair(khat):
        ref = 1.0 / T.ldr_alpha(eps, 0.15, khat, ehat)
        got = ldr_alpha_inv(torch.from_numpy(eps), 0.15, khat, ehat).numpy()
        assert np.allclose(got, ref, rtol=1e-12)


def test_ldr_alpha_inv_autograd_matches_closed_form():
    eps = torch.tensor([1.7 + 0.05j, 3.2 + 0.1j], dtype=torch.complex128, requires_grad=True)
    khat, ehat = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0], complex)
    kd = 0.151
    c = torch.tensor([0.3 - 0.7j, -1.1 + 0.2j], dtype=torch.complex128)
    loss = torch.real(torch.sum(c * ldr_alpha_inv(eps, kd, khat, ehat)))
    loss.backward()
    e = eps.detach().numpy()
    S = 0.0  # k along z, E along x
    dprime = -4 * math.pi / (e - 1) ** 2 + (T.LDR_B2 + T.LDR_B3 * S) * kd ** 2
    # real loss Re(c f(eps)): torch gradient = conj(c f'(eps))
    assert np.allclose(eps.grad.numpy(), np.conj(c.numpy() * dprime), rtol=1e-12)
