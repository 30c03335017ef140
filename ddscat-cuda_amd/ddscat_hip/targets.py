"""Synthetic targets and per-dipole operator setup for the standalone harness.

In DDSCAT these inputs come from the Fortran driver (TARGET/EXTEND, ALPHADIAG,
EVALE; ddscat_cuda_unified/fortran/getfml_cuda.f90:579-621), which sits above
the C-ABI boundary and is not part of this library.  They are restated here
from their published definitions so that tests and benchmarks can build the
same inputs the driver would hand to ``ddscat_cuda_solve_*_f32``:

* targets: sphere (sites with |r - c| <= D/2, c the lattice centre), block,
  seeded unions of ellipsoids;
* polarizability: lattice-dispersion relation of Draine & Goodman (1993,
  ApJ 405, 685; DDSCAT's LATTDR), alpha = a_CM / (1 + (a_CM/d^3)[(b1 + b2 m^2
  + b3 m^2 S)(kd)^2 - (2/3) i (kd)^3]), S = sum_j (a_j e_j)^2, b1=-1.8915316,
  b2=0.1648469, b3=-1.7700004; or the diagonal tensor of Gutkowicz-Krusin &
  Draine (2004; GKDLDR), whose jj element takes S_j = a_j^2; CXADIA = 1/alpha,
  CXAOFF = 0 for isotropic material;
* incident field: E_j = E0 exp(i k.r_j) at r_j = ((ix+1)+x0_1)dx_1, ...,
  the same site coordinates the far-field sums use
  (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:800-802).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

LDR_B1 = -1.8915316
LDR_B2 = 0.1648469
LDR_B3 = -1.7700004


@dataclass
class Target:
    """A dipole array on an NX x NY x NZ lattice; iocc[j] for j = ix + NX(iy + NY iz)."""
    nx: int
    ny: int
    nz: int
    iocc: np.ndarray                      # int16 [NAT]
    eps: np.ndarray | None = None         # complex128 [NAT0] per occupied dipole (None -> uniform)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nat(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def nat0(self) -> int:
        return int(np.count_nonzero(self.iocc))

    @property
    def occ(self) -> np.ndarray:
        return np.flatnonzero(self.iocc != 0).astype(np.int64)

    def centroid_x0(self) -> np.ndarray:
        """x0 such that the occupied-site centroid sits at the origin for the
        r_j = (i+1+x0) convention."""
        occ = self.occ
        ix = occ % self.nx
        iy = (occ // self.nx) % self.ny
        iz = occ // (self.nx * self.ny)
        return -np.array([ix.mean() + 1, iy.mean() + 1, iz.mean() + 1])

    def a_eff(self) -> float:
        """Effective radius in lattice units, (3 NAT0 / 4 pi)^(1/3)."""
        return (3.0 * self.nat0 / (4.0 * math.pi)) ** (1.0 / 3.0)


def sphere(diameter: int) -> Target:
    """Sites of a D^3 lattice within D/2 of the lattice centre ((D-1)/2 per axis)."""
    d = int(diameter)
    c = (d - 1) / 2.0
    i = np.arange(d)
    Z, Y, X = np.meshgrid(i, i, i, indexing="ij")
    r2 = (X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2
    iocc = (r2 <= (d / 2.0) ** 2).reshape(-1).astype(np.int16)
    return Target(d, d, d, iocc, name=f"sphere_D{d}")


def block(nx: int, ny: int | None = None, nz: int | None = None) -> Target:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return Target(nx, ny, nz, np.ones(nx * ny * nz, dtype=np.int16), name=f"block_{nx}x{ny}x{nz}")


def fft_friendly(n: int) -> int:
    """Smallest m >= n whose prime factors are 2, 3 and 5 (the lattice
    extension DDSCAT applies before its FFTs)."""
    m = max(1, int(n))
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def ellipsoid_union(rng: np.random.Generator, n_min=1, n_max=4, semi_lo=6.0, semi_hi=40.0,
                    grid_lo=24, grid_hi=96) -> Target:
    """Union of 1..4 randomly oriented ellipsoids with semi-axes U[6,40]
    (batched-solve workload generator).  The first ellipsoid sits at the
    centre; each further one is offset by up to half the first one's largest
    semi-axis per axis, clipped so that it stays inside a grid_hi-point box.
    Each axis of the bounding grid is at least grid_lo points and is extended
    to a 2,3,5-smooth size, so grids run from 24 to 96 points per axis."""
    half_max = grid_hi // 2 - 2
    k = int(rng.integers(n_min, n_max + 1))
    ells = []
    for e in range(k):
        ax = rng.uniform(semi_lo, semi_hi, size=3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        off = rng.uniform(-0.5, 0.5, size=3) * (ells[0][0].max() if e else 0.0)
        lim = max(0.0, half_max - ax.max())
        off = np.clip(off, -lim, lim)
        ells.append((ax, R, off))
    ext = np.max([np.abs(o) + a.max() for a, R, o in ells], axis=0)
    half = np.ceil(ext).astype(int) + 1
    dims = [fft_friendly(max(grid_lo, int(2 * h))) for h in half]
    nx, ny, nz = dims
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([ix - (nx - 1) / 2.0, iy - (ny - 1) / 2.0, iz - (nz - 1) / 2.0], axis=-1)
    occ = np.zeros(pts.shape[:-1], dtype=bool)
    for ax, R, off in ells:
        loc = (pts - off) @ R
        occ |= np.sum((loc / ax) ** 2, axis=-1) <= 1.0
    return Target(nx, ny, nz, occ.reshape(-1).astype(np.int16), name="ellipsoids",
                  meta={"n_ellipsoids": k})


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """n rotation matrices [n, 3, 3] uniform on SO(3) (normalised Gaussian
    quaternions).  Column 0 is the incidence direction in the target frame,
    columns 1 and 2 the two polarizations."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                     np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                     np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


def kd_for_size_parameter(target: Target, x: float) -> float:
    """k d for size parameter x = k a_eff."""
    return x / target.a_eff()


def ldr_s(khat, ehat, prescription="LATTDR") -> np.ndarray:
    """S of each component [3]: LATTDR sum_j (khat_j |ehat_j|)^2 (|ehat| = 1)
    for all three, GKDLDR khat_j^2."""
    khat = np.asarray(khat, float)
    khat = khat / np.linalg.norm(khat)
    if prescription == "GKDLDR":
        return khat ** 2
    if prescription != "LATTDR":
        raise ValueError(f"unknown polarizability prescription {prescription!r}")
    ea = np.abs(np.asarray(ehat, complex))
    S = float(np.sum((khat * ea / max(np.linalg.norm(ea), 1e-300)) ** 2))
    return np.full(3, S)


def ldr_alpha(eps, kd: float, khat, ehat, prescription="LATTDR") -> np.ndarray:
    """Lattice-dispersion-relation polarizability (units d^3): shape of eps
    for LATTDR (isotropic), [3, *eps.shape] (diagonal elements) for GKDLDR."""
    eps = np.asarray(eps, np.complex128)
    a_cm = (3.0 / (4.0 * math.pi)) * (eps - 1.0) / (eps + 2.0)
    S = ldr_s(khat, ehat, prescription)
    out = []
    for Sj in (S[:1] if prescription == "LATTDR" else S):
        corr = (LDR_B1 + LDR_B2 * eps + LDR_B3 * Sj * eps) * kd ** 2 - (2.0 / 3.0) * 1j * kd ** 3
        out.append(a_cm / (1.0 + a_cm * corr))
    return out[0] if prescription == "LATTDR" else np.stack(out)


def operators(target: Target, m_or_eps, kd: float, khat, ehat, use_eps=False, prescription="LATTDR"):
    """Full-layout CXADIA, CXAOFF [3*NAT] complex64 (zero on void sites)."""
    nat = target.nat
    occ = target.occ
    if target.eps is not None:
        eps = target.eps
    elif use_eps:
        eps = np.full(occ.size, complex(m_or_eps))
    else:
        eps = np.full(occ.size, complex(m_or_eps) ** 2)
    alpha = ldr_alpha(eps, kd, khat, ehat, prescription)
    adia = np.zeros((3, nat), dtype=np.complex128)
    adia[:, occ] = 1.0 / alpha
    aoff = np.zeros((3, nat), dtype=np.complex128)
    return adia.reshape(-1).astype(np.complex64), aoff.reshape(-1).astype(np.complex64)


def site_positions_full(target: Target, x0, dx=(1.0, 1.0, 1.0)) -> np.ndarray:
    j = np.arange(target.nat)
    ix = j % target.nx
    iy = (j // target.nx) % target.ny
    iz = j // (target.nx * target.ny)
    return np.stack([(ix + 1 + x0[0]) * dx[0], (iy + 1 + x0[1]) * dx[1], (iz + 1 + x0[2]) * dx[2]], axis=1)


def incident_field(target: Target, kvec, e0, x0, dx=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Full-layout E_inc [3*NAT] complex64: E0 exp(i k.r_j) on occupied sites, 0 elsewhere."""
    r = site_positions_full(target, x0, dx)
    ph = np.exp(1j * (r @ np.asarray(kvec, np.float64)))
    e = np.asarray(e0, np.complex128)[:, None] * ph[None, :]
    e[:, target.iocc == 0] = 0
    return e.reshape(-1).astype(np.complex64)


def polarization_pair(khat):
    """Two orthonormal real polarization vectors perpendicular to khat."""
    k = np.asarray(khat, float) / np.linalg.norm(khat)
    trial = np.array([0.0, 1.0, 0.0]) if abs(k[1]) < 0.9 else np.array([0.0, 0.0, 1.0])
    e1 = trial - k * np.dot(k, trial)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(k, e1)
    return e1.astype(np.complex128), e2.astype(np.complex128)
