"""Differentiable absorption efficiency for inverse design.

Q_abs = Cabs / (pi a_eff^2), with Cabs from EVALQ
(ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:939-950) of the solution of
A p = E_inc.  The forward pass is the batched device solve plus the batched
EVALQ; the backward pass is one batched adjoint solve on the same MATVEC
(ddscat_hip_cabs_adjoint: A is complex symmetric, so A^H lam = W p becomes
A mu = W conj(p)).  Everything runs through the HIP library; there is no
CPU path.

`AbsorptionEfficiency` maps per-RHS operators (adia, aoff) to Q_abs per RHS;
`ldr_alpha_inv` is the differentiable map from a permittivity map to the
lattice-dispersion-relation alpha^{-1}, so a design variable eps gets its
gradient by ordinary torch autograd.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import targets as T


def ldr_alpha_inv(eps: torch.Tensor, kd: float, khat, ehat) -> torch.Tensor:
    """1/alpha of the LDR polarizability (units d^-3) as a function of the
    complex permittivity tensor eps; the inverse of targets.ldr_alpha:
    1/alpha = 1/alpha_CM + (b1 + b2 eps + b3 S eps) kd^2 - (2/3) i kd^3."""
    khat = np.asarray(khat, float)
    khat = khat / np.linalg.norm(khat)
    ea = np.abs(np.asarray(ehat, complex))
    S = float(np.sum((khat * ea / max(np.linalg.norm(ea), 1e-300)) ** 2))
    inv_cm = (4.0 * math.pi / 3.0) * (eps + 2.0) / (eps - 1.0)
    corr = (T.LDR_B1 + (T.LDR_B2 + T.LDR_B3 * S) * eps) * kd ** 2 - (2.0 / 3.0) * 1j * kd ** 3
    return inv_cm + corr


class NotConverged(RuntimeError):
    """A forward or adjoint solve stopped at its iteration limit above the tolerance."""


def _require_converged(which, tole, tol, maxit):
    """The batched solvers return 0 at the iteration limit with the achieved
    residual (the reference's convention); a Q_abs or gradient from such a
    solve must not reach the caller silently."""
    tole = np.asarray(tole, np.float64)
    bad = np.flatnonzero(~(tole <= 1.01 * tol))
    if bad.size:
        raise NotConverged(f"{which} solve: RHS {bad.tolist()} at residual {tole[bad].tolist()} > tol {tol} "
                           f"after {maxit} iterations")


class AbsorptionEfficiency(torch.autograd.Function):
    """Q_abs[r] for RHS r of the prepared target.

    Inputs: adia, aoff complex64 CUDA tensors [B, 3*NAT0] (compact, per RHS),
    e (incident field, same shape, not differentiated), lib (ddscat_hip
    Library with the target prepared), ak [B, 3] (k d per RHS), geo = pi
    a_eff^2 in d^2, solver name, tolerance, iteration limit.  Output:
    float64 CUDA tensor [B]."""

    @staticmethod
    def forward(ctx, adia, aoff, e, lib, ak, geo, method="qmrccg", tol=1e-5, maxit=500):
        assert adia.is_cuda and adia.dtype == torch.complex64 and adia.is_contiguous()
        assert aoff.shape == adia.shape and e.shape == adia.shape
        B = adia.shape[0]
        aoff = aoff.contiguous()
        e = e.contiguous()
        x = torch.zeros_like(e)
        torch.cuda.synchronize()
        _, tole, _, _ = lib.solve_batch(method, B, e.data_ptr(), x.data_ptr(), adia.data_ptr(), aoff.data_ptr(),
                                        op_per_rhs=True, maxit=maxit, tol=tol)
        _require_converged("forward", tole, tol, maxit)
        q = lib.evalq_batch(B, e.data_ptr(), x.data_ptr(), adia.data_ptr(), aoff.data_ptr(), ak)
        ctx.save_for_backward(adia, aoff, x)
        ctx.lib, ctx.ak, ctx.geo, ctx.method, ctx.tol, ctx.maxit = lib, ak, geo, method, tol, maxit
        return torch.as_tensor(q[:, 0] / geo, dtype=torch.float64, device=adia.device)

    @staticmethod
    def backward(ctx, gq):
        adia, aoff, x = ctx.saved_tensors
        B = adia.shape[0]
        g_adia = torch.empty_like(adia)
        g_aoff = torch.empty_like(aoff)
        torch.cuda.synchronize()
        _, tole, _ = ctx.lib.cabs_adjoint(ctx.method, B, x.data_ptr(), adia.data_ptr(), aoff.data_ptr(), ctx.ak,
                                          g_adia.data_ptr(), g_aoff.data_ptr(), maxit=ctx.maxit, tol=ctx.tol)
        _require_converged("adjoint", tole, ctx.tol, ctx.maxit)
        # real loss L: dL = Re(g . dz)  ->  torch's gradient for z is conj(g)
        w = (gq.to(torch.float32) / ctx.geo).reshape(B, 1)
        return (torch.conj(g_adia) * w, torch.conj(g_aoff) * w, None, None, None, None, None, None, None)


def design_q_abs(lib, target, eps: torch.Tensor, kd: float, khat, pols, method="qmrccg", tol=1e-5, maxit=500):
    """Mean Q_abs over the polarizations `pols` for incidence khat, as a
    differentiable function of the per-dipole permittivity eps (complex
    tensor [NAT0], any device).  The target must be prepared in `lib`."""
    occ = target.occ
    khat = np.asarray(khat, float) / np.linalg.norm(khat)
    dev = torch.device("cuda")
    rows_ad, rows_e = [], []
    for e0 in pols:
        ai = ldr_alpha_inv(eps.to(dev), kd, khat, e0).to(torch.complex64)
        rows_ad.append(ai.repeat(3))
        ef = T.incident_field(target, khat * kd, e0, target.centroid_x0())
        ec = np.ascontiguousarray(ef.reshape(3, -1)[:, occ].reshape(-1))
        rows_e.append(torch.from_numpy(ec).to(dev))
    adia = torch.stack(rows_ad).contiguous()
    aoff = torch.zeros_like(adia)
    e = torch.stack(rows_e).contiguous()
    geo = math.pi * (3.0 * target.nat0 / (4.0 * math.pi)) ** (2.0 / 3.0)
    ak = np.tile((khat * kd).astype(np.float32), (len(pols), 1))
    assert lib.nat0 == target.nat0, "prepare the target in the library first"
    q = AbsorptionEfficiency.apply(adia, aoff, e, lib, ak, geo, method, tol, maxit)
    return q.mean()
