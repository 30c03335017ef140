"""FP64 numpy restatement of the DDSCAT Green-tensor MATVEC and far-field sums.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.

Every function names the reference lines it restates (paths relative to the
reference repository root, ``cu`` = ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu,
``cp`` = ddscat_cuda_unified/fortran/cprod_cuda.f90).

Pinning (what fixes each function independently of its own formula):
  * green_elements (IDIPINT=0): an independent restatement of the radiating
    point-dipole field E = e^{ikr}[k^2 (n x p) x n / r + (3n(n.p) - p)(1/r^3 - ik/r^2)]
    (test_oracle_green.py) and the FCD form with its filter terms switched off.
  * cisi: scipy.special.sici.
  * green_octant: brute-force spectral sums on tiny grids and the dense-matrix
    MATVEC below (test_oracle_matvec.py).
  * green_sum_lag / green_full_pbc (IPBC=1): Bloch periodicity
    G(l + P) = e^{i k.P} G(l), the replica window against a much wider
    brute-force replica sum, the undamped isolated term on the stationary-phase
    line, and the dense operator of replica sums against the full-grid FFT
    MATVEC (test_oracle_pbc.py).
  * matvec_fft: dense 3NAT x 3NAT matrix built directly from green_elements on
    2^3..6^3 grids with holes; complex symmetry A^T = A; 'C' identity.
  * evalq / scat: the reference's own numpy identities
    (scripts/test_gpu_final_sections_math.py, restated as own tests), a single
    dipole's closed-form radiated power, the optical theorem for converged
    solutions, and Mie theory for a sphere at DDA accuracy.
"""
from __future__ import annotations

import math
import numpy as np

EULER_GAMMA = 0.57721566490153286061
PI = math.pi

# component order of the symmetric 3x3 tensor: xx, xy, xz, yy, yz, zz  (cu:625-626)
COMP_IJ = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
# parity (x, y, z) of each component in the lag vector: 0 even, 1 odd  (cu:112)
COMP_PARITY = ((0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 0, 0), (0, 1, 1), (0, 0, 0))


# ----------------------------------------------------------------------------
# Ci / Si   (cu:610-621 restated from the mathematics: power series for
# |x| <= 2, Lentz continued fraction of E1(ix) otherwise)
# ----------------------------------------------------------------------------
def cisi(x: float) -> tuple[float, float]:
    """Cosine and sine integrals Ci(|x|), sign(x) Si(|x|)  (cu:610-621)."""
    t = abs(float(x))
    if t == 0.0:
        return -1.0e30, 0.0
    if t > 2.0:
        # e^{z} E1(z) = 1/(z+1 - 1/(z+3 - 4/(z+5 - 9/(...)))),  z = i t
        z = complex(0.0, t)
        tiny = 1e-300
        b = z + 1.0
        c = 1.0 / tiny
        d = 1.0 / b
        h = d
        for n in range(2, 200):
            a = -float((n - 1) * (n - 1))
            b = b + 2.0
            d = 1.0 / (a * d + b)
            c = b + a / c
            dl = c * d
            h = h * dl
            if abs(dl.real - 1.0) + abs(dl.imag) < 1e-16:
                break
        e1 = h * complex(math.cos(t), -math.sin(t))      # E1(i t)
        ci = -e1.real
        si = PI / 2 + e1.imag
    else:
        si = 0.0
        term = t
        k = 0
        while True:
            val = term / (2 * k + 1)
            si += val
            k += 1
            term *= -t * t / ((2 * k) * (2 * k + 1))
            if abs(val) < 1e-18 * max(abs(si), 1e-300) or k > 60:
                break
        cs = 0.0
        term = -t * t / 2.0
        k = 1
        while True:
            val = term / (2 * k)
            cs += val
            term *= -t * t / ((2 * k + 1) * (2 * k + 2))
            k += 1
            if abs(val) < 1e-18 * max(abs(cs), 1e-300) or k > 60:
                break
        ci = EULER_GAMMA + math.log(t) + cs
    if x < 0:
        si = -si
    return ci, si


# ----------------------------------------------------------------------------
# Green-tensor elements  (cu:622-633)
# ----------------------------------------------------------------------------
def green_elements(x, y, z, akd: float, idipint: int = 0) -> np.ndarray:
    """Six components (xx,xy,xz,yy,yz,zz) of the direct interaction tensor c
    for displacement (x,y,z) in lattice units (cu:622-633, IPBC=0 so the PBC
    phase and damping are zero).  Entries with r^2 <= 1e-6 are zero.

    IDIPINT=0: c_ij = -e^{ikr}/r^3 [k^2 (x_i x_j - r^2 d_ij) + (1-ikr)/r^2 (r^2 d_ij - 3 x_i x_j)]
    which is the field propagator G so that E_j = sum_l G(r_j - r_l) P_l.
    IDIPINT=1: filtered coupled dipole (cu:627-631)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    shape = np.broadcast(x, y, z).shape
    if shape == ():
        return green_elements(np.atleast_1d(x), np.atleast_1d(y), np.atleast_1d(z), akd, idipint)[:, 0]
    x, y, z = (np.broadcast_to(a, shape).astype(np.float64) for a in (x, y, z))
    out = np.zeros((6,) + shape, dtype=np.complex128)
    r2 = x * x + y * y + z * z
    m = r2 > 1e-6
    if not np.any(m):
        return out
    xm, ym, zm, r2m = x[m], y[m], z[m], r2[m]
    r = np.sqrt(r2m)
    r3 = r * r2m
    kr = akd * r
    xx = (xm, ym, zm)
    akd2 = akd * akd
    if idipint == 0:
        ph = np.exp(1j * kr) / r3
        fac = (1.0 - 1j * kr) / r2m
        for c, (i, j) in enumerate(COMP_IJ):
            if i == j:
                q = xx[i] * xx[i]
                term = akd2 * (q - r2m) + fac * (r2m - 3.0 * q)
                out[c][m] = -ph * term
            else:
                term2 = akd2 - 3.0 * fac
                out[c][m] = -ph * term2 * xx[i] * xx[j]
        return out
    if idipint != 1:
        raise ValueError("IDIPINT must be 0 or 1")
    kf = PI
    kfr = kf * r
    cip = np.empty_like(r); sip = np.empty_like(r)
    cim = np.empty_like(r); sim = np.empty_like(r)
    for n in range(r.size):
        cip[n], sip[n] = cisi(kfr[n] + kr[n])
        cim[n], sim[n] = cisi(kfr[n] - kr[n])
    sinkr, coskr = np.sin(kr), np.cos(kr)
    sinkfr, coskfr = np.sin(kfr), np.cos(kfr)
    a0 = sinkr * (cip - cim) + coskr * (PI - sip - sim)
    a1 = akd * sinkr * (-PI + sip + sim) + akd * coskr * (cip - cim) - 2.0 * sinkfr / r
    a2 = akd2 * (sinkr * (cim - cip) + coskr * (sip + sim - PI)) + 2.0 * (sinkfr - kfr * coskfr) / r2m
    return _fcd_assemble(out, m, xx, r, r2m, r3, kr, akd2, kfr, sinkfr, coskfr, a0, a1, a2)


def _fcd_assemble(out, m, xx, r, r2m, r3, kr, akd2, kfr, sinkfr, coskfr, a0, a1, a2):
    eikr = np.exp(1j * kr)
    g0 = (eikr - a0 / PI) / r
    g1 = (eikr * (-1.0 + 1j * kr) + (a0 - a1 * r) / PI) / r2m
    g2 = (eikr * (2.0 - kr * kr - 2j * kr) - (2.0 * (a0 - r * a1) + r2m * a2) / PI) / r3
    h2 = (sinkfr - kfr * coskfr) / (PI * r3)
    radial = g2 / r2m - g1 / r3
    for c, (i, j) in enumerate(COMP_IJ):
        if i == j:
            base = akd2 * g0 + g1 / r + (2.0 / 3.0) * h2 + radial * xx[i] * xx[i]
            out[c][m] = base
        else:
            out[c][m] = radial * xx[i] * xx[j]
    return out


def fcd_without_filter(x, y, z, akd):
    """FCD assembly (cu:629-631) with a0=a1=a2=h2=0: must equal IDIPINT=0."""
    x = np.atleast_1d(np.asarray(x, float)); y = np.atleast_1d(np.asarray(y, float))
    z = np.atleast_1d(np.asarray(z, float))
    out = np.zeros((6,) + x.shape, dtype=np.complex128)
    r2 = x * x + y * y + z * z
    m = r2 > 1e-6
    r = np.sqrt(r2[m]); r3 = r * r2[m]; kr = akd * r
    zero = np.zeros_like(r)
    kfr = np.zeros_like(r)
    return _fcd_assemble(out, m, (x[m], y[m], z[m]), r, r2[m], r3, kr, akd * akd,
                         kfr, zero, np.ones_like(r), zero, zero, zero)


# ----------------------------------------------------------------------------
# Spectral Green tensor  (cu:644-652, IPBC=0)
# ----------------------------------------------------------------------------
def lag_axis(n: int) -> np.ndarray:
    """Lag for padded index 0..2n-1: l = i (i<n), i-2n (i>n), Nyquist i=n -> NaN
    marker (the reference zeroes that plane, cu:644)."""
    i = np.arange(2 * n)
    lag = np.where(i < n, i, i - 2 * n).astype(np.float64)
    return lag


def green_lag_grid(nx, ny, nz, akd, dx=(1.0, 1.0, 1.0), idipint=0) -> np.ndarray:
    """Full lag grid [6][2nz][2ny][2nx] (cu:644): FP64 Green at lag (lx,ly,lz)*dx,
    Nyquist planes (ix==nx, iy==ny, iz==nz) zero."""
    lx, ly, lz = lag_axis(nx), lag_axis(ny), lag_axis(nz)
    Z, Y, X = np.meshgrid(lz * dx[2], ly * dx[1], lx * dx[0], indexing="ij")
    g = green_elements(X, Y, Z, akd, idipint)
    g[:, nz, :, :] = 0
    g[:, :, ny, :] = 0
    g[:, :, :, nx] = 0
    return g


def green_octant(nx, ny, nz, akd, dx=(1.0, 1.0, 1.0), idipint=0) -> np.ndarray:
    """Trimmed spectral Green [6][nz+1][ny+1][nx+1] (cu:647-652, IPBC=0):
    fill the lag grid, 3-D transform with the +i exponent (CUFFT_INVERSE,
    unnormalised), keep the (nx+1)(ny+1)(nz+1) octant.  Memory order matches
    the reference's b = ir + (nx+1)(jr + (ny+1) kr) (cu:112) and the Fortran
    CXZC(NX+1,NY+1,NZ+1,6) array of the CPU path (cp:308)."""
    g = green_lag_grid(nx, ny, nz, akd, dx, idipint)
    G = (2 * nx) * (2 * ny) * (2 * nz)
    gh = np.fft.ifftn(g, axes=(1, 2, 3)) * G
    return np.ascontiguousarray(gh[:, : nz + 1, : ny + 1, : nx + 1])


def green_full_from_octant(goct: np.ndarray, nx, ny, nz) -> np.ndarray:
    """Rebuild the full spectral tensor [6][2nz][2ny][2nx] from the octant with
    the parity signs of green_kernel (cu:109-115)."""
    ix = np.arange(2 * nx); iy = np.arange(2 * ny); iz = np.arange(2 * nz)
    ir = np.minimum(ix, 2 * nx - ix); jr = np.minimum(iy, 2 * ny - iy); kr = np.minimum(iz, 2 * nz - iz)
    sx = np.where(ix <= nx, 1.0, -1.0); sy = np.where(iy <= ny, 1.0, -1.0); sz = np.where(iz <= nz, 1.0, -1.0)
    full = goct[:, kr][:, :, jr][:, :, :, ir]
    s3 = (sx[None, None, :], sy[None, :, None], sz[:, None, None])
    for c, par in enumerate(COMP_PARITY):
        sgn = np.ones((2 * nz, 2 * ny, 2 * nx))
        for a in range(3):
            if par[a]:
                sgn = sgn * s3[a]
        full[c] = full[c] * sgn
    return full


# ----------------------------------------------------------------------------
# Periodic targets, IPBC=1  (cu:634-645, cu:650)
# ----------------------------------------------------------------------------
def pbc_window(x0, y0, z0, akvec, gamma, pyddx, pzddx):
    """Replica window of one lag (cu:636-640): (jpbc, jpy1, jpy2, jpz1, jpz2,
    krf, phasf).  jpbc 1: periodic in y only (pzddx <= 0), 2: in z only
    (pyddx <= 0), 3: both.  Bounds truncate toward zero like C's (int)."""
    akx, aky, akz = (float(v) for v in akvec)
    akd = math.sqrt(akx * akx + aky * aky + akz * akz)
    if pyddx <= 0.0:
        jpbc, kp2 = 2, akx * akx + aky * aky
    elif pzddx <= 0.0:
        jpbc, kp2 = 1, akx * akx + akz * akz
    else:
        jpbc, kp2 = 3, akx * akx
    if kp2 <= 0.0:
        raise ValueError("k has no component across the periodic axes (status 13)")
    kperp = math.sqrt(kp2)
    phimax = 1.0 / gamma
    term0 = akd * akd * phimax * phimax / kp2
    term1 = 2.0 * akd * phimax
    term2y = aky * phimax / kp2
    term2z = akz * phimax / kp2
    rperp = (math.sqrt(x0 * x0 + z0 * z0) if jpbc == 1 else
             math.sqrt(x0 * x0 + y0 * y0) if jpbc == 2 else abs(x0))
    term3 = math.sqrt(term0 + term1 * rperp) / kperp
    krf = akd * akd * rperp / kperp
    jpy1 = jpy2 = jpz1 = jpz2 = 0
    yf = zf = 0.0
    if jpbc != 2:
        yf = y0 - aky * rperp / kperp
        jpy1 = int((yf - term2y - term3) / pyddx + 0.999999)
        jpy2 = int((yf - term2y + term3) / pyddx)
    if jpbc != 1:
        zf = z0 - akz * rperp / kperp
        jpz1 = int((zf - term2z - term3) / pzddx + 0.999999)
        jpz2 = int((zf - term2z + term3) / pzddx)
    phasf = aky * yf if jpbc == 1 else akz * zf if jpbc == 2 else aky * yf + akz * zf
    return jpbc, jpy1, jpy2, jpz1, jpz2, krf, phasf


def green_sum_lag(lx, ly, lz, akvec, gamma, pyd, pzd, dx=(1.0, 1.0, 1.0), idipint=0, window=None):
    """Six components at lag (lx, ly, lz) of a periodic target (cu:634-643):
    sum over source replicas J = (jpy, jpz), displaced by J*(PYD*DX2*DX2,
    PZD*DX3*DX3) as the reference displaces them, of
        G(r_lag - J P) e^{i k.J P} e^{-beta phi^4},  beta = 16 gamma^4,
        phi = k|r| - krf + k.J P - phasf.
    ``window`` = (jpy1, jpy2, jpz1, jpz2) replaces the computed window (tests)."""
    akvec = tuple(float(v) for v in akvec)
    akd = math.sqrt(sum(v * v for v in akvec))
    x0, y0, z0 = lx * dx[0], ly * dx[1], lz * dx[2]
    pyddx, pzddx = pyd * dx[1], pzd * dx[2]
    jpbc, jpy1, jpy2, jpz1, jpz2, krf, phasf = pbc_window(x0, y0, z0, akvec, gamma, pyddx, pzddx)
    if window is not None:
        jpy1, jpy2, jpz1, jpz2 = window
    if jpy2 < jpy1 or jpz2 < jpz1:
        return np.zeros(6, np.complex128)
    JY, JZ = np.meshgrid(np.arange(jpy1, jpy2 + 1, dtype=np.float64),
                         np.arange(jpz1, jpz2 + 1, dtype=np.float64), indexing="ij")
    rjpy = JY * pyddx * dx[1]
    rjpz = JZ * pzddx * dx[2]
    y, z = y0 - rjpy, z0 - rjpz
    kr = akd * np.sqrt(x0 * x0 + y * y + z * z)
    phasyz = akvec[1] * rjpy + akvec[2] * rjpz
    beta = 16.0 * gamma ** 4
    phi = kr - krf + phasyz - phasf
    f = np.exp(-beta * phi ** 4) * np.exp(1j * phasyz)
    g = green_elements(np.full_like(y, x0), y, z, akd, idipint)
    return (g * f[None]).reshape(6, -1).sum(axis=1)


def green_lag_grid_pbc(nx, ny, nz, akvec, gamma, pyd, pzd, dx=(1.0, 1.0, 1.0), idipint=0) -> np.ndarray:
    """Whole lag grid [6][2nz][2ny][2nx] of replica sums, Nyquist planes zero
    (cu:645)."""
    out = np.zeros((6, 2 * nz, 2 * ny, 2 * nx), np.complex128)
    lag = lambda i, n: i if i < n else i - 2 * n
    for iz in range(2 * nz):
        for iy in range(2 * ny):
            for ix in range(2 * nx):
                if ix == nx or iy == ny or iz == nz:
                    continue
                out[:, iz, iy, ix] = green_sum_lag(lag(ix, nx), lag(iy, ny), lag(iz, nz), akvec, gamma, pyd,
                                                   pzd, dx, idipint)
    return out


def green_full_pbc(nx, ny, nz, akvec, gamma, pyd, pzd, dx=(1.0, 1.0, 1.0), idipint=0) -> np.ndarray:
    """Whole spectral Green [6][2nz][2ny][2nx] for IPBC=1 (cu:650): the lag
    grid transformed with the +i exponent, unnormalised, nothing trimmed.
    This is the CXZC array ddscat_cuda_prepare_f32 takes for IPBC=1 (cu:1020)."""
    g = green_lag_grid_pbc(nx, ny, nz, akvec, gamma, pyd, pzd, dx, idipint)
    G = (2 * nx) * (2 * ny) * (2 * nz)
    return np.fft.ifftn(g, axes=(1, 2, 3)) * G


def dense_green_matrix_pbc(nx, ny, nz, akvec, gamma, pyd, pzd, dx=(1.0, 1.0, 1.0), idipint=0) -> np.ndarray:
    """Dense interaction matrix of a periodic target, Gm[c*NAT+j, d*NAT+l] =
    G_pbc(r_j - r_l)_cd from green_sum_lag directly (no FFT).  G_pbc is not
    even in the lag, so the orientation r_j - r_l matters."""
    nat = nx * ny * nz
    cache = {}
    M = np.zeros((3 * nat, 3 * nat), np.complex128)
    j = np.arange(nat)
    ix, iy, iz = j % nx, (j // nx) % ny, j // (nx * ny)
    for a in range(nat):
        for b in range(nat):
            key = (ix[a] - ix[b], iy[a] - iy[b], iz[a] - iz[b])
            if key not in cache:
                cache[key] = green_sum_lag(*key, akvec, gamma, pyd, pzd, dx, idipint)
            g = cache[key]
            for c, (p, q) in enumerate(COMP_IJ):
                M[p * nat + a, q * nat + b] = g[c]
                M[q * nat + a, p * nat + b] = g[c]
    return M


# ----------------------------------------------------------------------------
# MATVEC  (cu:104-118 FFT3D; CPU semantics cp:446-607)
# ----------------------------------------------------------------------------
def _finish(cwhat, x3, adia3, aoff3, w3, iocc):
    """y = adia*x - w + offdiag (cu:116-118, cp:503-515), mask and 'C'."""
    a23, a31, a12 = aoff3[0], aoff3[1], aoff3[2]
    x1, x2, xz = x3[0], x3[1], x3[2]
    y = adia3 * x3 - w3
    y[0] += a31 * xz + a12 * x2
    y[1] += a12 * x1 + a23 * xz
    y[2] += a23 * x2 + a31 * x1
    if cwhat == "C":
        y = np.conj(y)
    if iocc is not None:
        y[:, np.asarray(iocc).reshape(-1) == 0] = 0
    return y


def matvec_fft(cwhat, x, adia, aoff, goct, nx, ny, nz, iocc=None, dtype=np.complex128,
               workers=None, ipbc=0):
    """y = A x with A = adia + offdiag - G (cu:406-415 + kernels cu:104-118).
    goct: the trimmed octant (IPBC=0, parity reconstruction) or, with ipbc=1,
    the whole spectral grid [6][2nz][2ny][2nx] used as it is (cu:113).
    x, adia, aoff: full component-major [3*NAT] arrays, site j = ix + nx(iy + ny iz).
    'C' applies conj(A conj(x)) (cp:519-593).  dtype complex64 is the FP32
    CPU baseline (scipy.fft with ``workers`` threads); complex128 the oracle."""
    nat = nx * ny * nz
    cwhat = chr(cwhat) if isinstance(cwhat, int) else cwhat
    x3 = np.asarray(x).reshape(3, nat).astype(dtype)
    if cwhat == "C":
        x3 = np.conj(x3)
    G = 8 * nat
    w = np.zeros((3, 2 * nz, 2 * ny, 2 * nx), dtype=dtype)
    w[:, :nz, :ny, :nx] = x3.reshape(3, nz, ny, nx)
    if workers is None and dtype == np.complex128:
        wh = np.fft.ifftn(w, axes=(1, 2, 3)) * G
    else:
        import scipy.fft as sfft
        wh = sfft.ifftn(w, axes=(1, 2, 3), workers=workers, overwrite_x=True) * dtype(G).real
    gf = (np.asarray(goct) if ipbc else green_full_from_octant(goct, nx, ny, nz)).astype(dtype)
    p0, p1, p2 = wh[0], wh[1], wh[2]
    q = np.empty_like(wh)
    q[0] = gf[0] * p0 + gf[1] * p1 + gf[2] * p2
    q[1] = gf[1] * p0 + gf[3] * p1 + gf[4] * p2
    q[2] = gf[2] * p0 + gf[4] * p1 + gf[5] * p2
    del wh
    if workers is None and dtype == np.complex128:
        wb = np.fft.fftn(q, axes=(1, 2, 3))
    else:
        import scipy.fft as sfft
        wb = sfft.fftn(q, axes=(1, 2, 3), workers=workers, overwrite_x=True)
    w3 = wb[:, :nz, :ny, :nx].reshape(3, nat) / dtype(G).real
    y = _finish(cwhat, x3, np.asarray(adia).reshape(3, nat).astype(dtype),
                np.asarray(aoff).reshape(3, nat).astype(dtype), w3, iocc)
    return y.reshape(-1)


class CpuMatvecFP32:
    """The FP32 CPU MATVEC used as the timed CPU baseline: the same operator as
    matvec_fft in complex64, with the spectral Green tensor expanded to the
    full padded grid once (as the CPU path keeps CXZC between calls,
    cp:308) and the 3-D transforms run by scipy.fft on ``workers`` threads."""

    def __init__(self, goct, nx, ny, nz, workers=None):
        import scipy.fft as sfft
        self.sfft = sfft
        self.n = (nx, ny, nz)
        self.workers = workers
        self.gf = green_full_from_octant(goct, nx, ny, nz).astype(np.complex64)
        self.w = np.zeros((3, 2 * nz, 2 * ny, 2 * nx), dtype=np.complex64)

    def apply(self, cwhat, x, adia, aoff, iocc=None):
        nx, ny, nz = self.n
        nat = nx * ny * nz
        G = np.float32(8 * nat)
        x3 = np.asarray(x, np.complex64).reshape(3, nat)
        if cwhat == "C":
            x3 = np.conj(x3)
        w = self.w
        w[...] = 0
        w[:, :nz, :ny, :nx] = x3.reshape(3, nz, ny, nx)
        wh =self.sfft.ifftn(w, axes=(1, 2, 3), workers=self.workers, overwrite_x=True, norm="forward")
        g = self.gf
        q = np.empty_like(wh)
        np.multiply(g[0], wh[0], out=q[0]); q[0] += g[1] * wh[1]; q[0] += g[2] * wh[2]
        np.multiply(g[1], wh[0], out=q[1]); q[1] += g[3] * wh[1]; q[1] += g[4] * wh[2]
        np.multiply(g[2], wh[0], out=q[2]); q[2] += g[4] * wh[1]; q[2] += g[5] * wh[2]
        del wh
        wb = self.sfft.fftn(q, axes=(1, 2, 3), workers=self.workers, overwrite_x=True)
        w3 = wb[:, :nz, :ny, :nx].reshape(3, nat) / G
        y = _finish(cwhat, x3, np.asarray(adia, np.complex64).reshape(3, nat),
                    np.asarray(aoff, np.complex64).reshape(3, nat), w3, iocc)
        return y.reshape(-1)


def green_octant_separable(nx, ny, nz, akd, dx=(1.0, 1.0, 1.0), idipint=0) -> np.ndarray:
    """The trimmed spectral octant of green_octant (cu:647-652) from (N+1)^3
    closed-form evaluations: each component has a definite parity per axis
    (COMP_PARITY), so the +i transform of the lag grid restricted to the
    octant is a cos sum (even) or i times a sin sum (odd) per axis.  Used for
    the large-grid oracle runs, where the full FP64 lag grid and its 3-D FFT
    would be slow; equal to green_octant to 1e-13 (test_oracle_green.py)."""
    ls = [np.arange(n + 1) for n in (nx, ny, nz)]
    Z, Y, X = np.meshgrid(ls[2] * dx[2], ls[1] * dx[1], ls[0] * dx[0], indexing="ij")
    g = green_elements(X, Y, Z, akd, idipint)
    g[:, nz] = 0
    g[:, :, ny] = 0
    g[:, :, :, nx] = 0

    def tmat(n, odd):
        k = np.arange(n + 1)[:, None]
        l = np.arange(n + 1)[None, :]
        m = 2 * (np.sin if odd else np.cos)(np.pi * k * l / n)
        m[:, 0] = 0 if odd else 1
        m[:, n] = 0
        return m

    out = np.empty_like(g)
    for c, par in enumerate(COMP_PARITY):
        a = np.einsum("kl,zyl->zyk", tmat(nx, par[0]), g[c], optimize=True)
        a = np.einsum("kl,zlx->zkx", tmat(ny, par[1]), a, optimize=True)
        a = np.einsum("kl,lyx->kyx", tmat(nz, par[2]), a, optimize=True)
        # two odd axes: i * i = -1 (xy, xz, yz)
        out[c] = -a if sum(par) == 2 else a
    return out


class FftMatvec64:
    """FP64 FFT MATVEC of matvec_fft on compact [3][NAT0] vectors, for the
    large-grid oracle runs (128^3 and 64^3 blocks, ellipsoid unions): the same
    operator (cu:104-118, cp:446-607), with the spectral Green expanded to the
    full grid once and the zero-padded transforms pruned axis by axis (x on
    the NZ*NY input rows, y on the NZ input planes, z on everything; the
    inverse keeps iz < NZ, then iy < NY, then ix < NX).  scipy.fft in
    complex128 on ``workers`` threads.  Equal to matvec_fft at 1e-13
    (test_oracle_matvec.py)."""

    def __init__(self, goct, nx, ny, nz, iocc, workers=None, ipbc=0):
        import scipy.fft as sfft
        self.sfft = sfft
        self.n = (nx, ny, nz)
        self.workers = workers
        self.occ = occupied_index(iocc)
        self.gf = np.asarray(goct, np.complex128) if ipbc else green_full_from_octant(goct, nx, ny, nz)

    def full(self, cwhat, x3, adia3, aoff3):
        """y [3][NAT] from full-layout x, adia, aoff [3][NAT] (no IOCC mask)."""
        nx, ny, nz = self.n
        f = self.sfft
        w = self.workers
        xs = np.conj(x3) if cwhat == "C" else np.asarray(x3, np.complex128)
        a = np.zeros((3, nz, ny, 2 * nx), np.complex128)
        a[..., :nx] = xs.reshape(3, nz, ny, nx)
        a = f.ifft(a, axis=3, norm="forward", workers=w, overwrite_x=True)
        b = np.zeros((3, nz, 2 * ny, 2 * nx), np.complex128)
        b[:, :, :ny] = a
        del a
        b = f.ifft(b, axis=2, norm="forward", workers=w, overwrite_x=True)
        c = np.zeros((3, 2 * nz, 2 * ny, 2 * nx), np.complex128)
        c[:, :nz] = b
        del b
        c = f.ifft(c, axis=1, norm="forward", workers=w, overwrite_x=True)
        q = self._green_multiply(c)
        del c
        q = f.fft(q, axis=1, workers=w, overwrite_x=True)[:, :nz]
        q = f.fft(q, axis=2, workers=w, overwrite_x=True)[:, :, :ny]
        q = f.fft(q, axis=3, workers=w, overwrite_x=True)[..., :nx]
        w3 = q.reshape(3, nx * ny * nz) / float(8 * nx * ny * nz)
        y = _finish(cwhat, xs, np.asarray(adia3, np.complex128), np.asarray(aoff3, np.complex128), w3, None)
        return y

    def _green_multiply(self, c):
        """q = G c per spectral point (symmetric 3x3), in slabs of kz planes
        spread over ``workers`` threads (numpy releases the GIL)."""
        from concurrent.futures import ThreadPoolExecutor
        g = self.gf
        q = np.empty_like(c)
        nzz = c.shape[1]
        step = max(1, nzz // (4 * (self.workers or 1)))

        def slab(z0):
            z = slice(z0, min(nzz, z0 + step))
            c0, c1, c2 = c[0, z], c[1, z], c[2, z]
            for i, (a, b, d) in enumerate(((0, 1, 2), (1, 3, 4), (2, 4, 5))):
                t = q[i, z]
                np.multiply(g[a, z], c0, out=t)
                t += g[b, z] * c1
                t += g[d, z] * c2

        with ThreadPoolExecutor(max_workers=self.workers or 1) as ex:
            list(ex.map(slab, range(0, nzz, step)))
        return q

    def __call__(self, cwhat, xc, adia_c, aoff_c):
        """y [3*NAT0] = A x on compact vectors with compact operators."""
        nx, ny, nz = self.n
        nat = nx * ny * nz
        occ = self.occ
        n0 = occ.size
        full = lambda v: scatter_compact(np.asarray(v, np.complex128), occ, nat).reshape(3, nat)
        y = self.full(cwhat, full(xc), full(adia_c), full(aoff_c))
        return np.ascontiguousarray(y[:, occ]).reshape(3 * n0)


def dense_green_matrix(nx, ny, nz, akd, dx=(1.0, 1.0, 1.0), idipint=0) -> np.ndarray:
    """Dense interaction matrix Gm[3*NAT, 3*NAT] (component-major ordering),
    Gm[c*NAT+j, d*NAT+l] = c_cd(r_j - r_l), zero on the diagonal blocks j==l.
    Built straight from green_elements with no FFT: the brute-force pin."""
    nat = nx * ny * nz
    j = np.arange(nat)
    ix, iy, iz = j % nx, (j // nx) % ny, j // (nx * ny)
    DX = (ix[:, None] - ix[None, :]) * dx[0]
    DY = (iy[:, None] - iy[None, :]) * dx[1]
    DZ = (iz[:, None] - iz[None, :]) * dx[2]
    g = green_elements(DX, DY, DZ, akd, idipint)
    M = np.zeros((3 * nat, 3 * nat), dtype=np.complex128)
    for c, (a, b) in enumerate(COMP_IJ):
        M[a * nat:(a + 1) * nat, b * nat:(b + 1) * nat] = g[c]
        if a != b:
            M[b * nat:(b + 1) * nat, a * nat:(a + 1) * nat] = g[c]
    return M


def dense_operator(nx, ny, nz, akd, adia, aoff, iocc=None, dx=(1.0, 1.0, 1.0), idipint=0, gm=None):
    """Full dense A = diag(adia) + offdiag(aoff) - Gm (cp:503-515), rows and
    columns of unoccupied sites zeroed (the MATVEC masks y there and x is
    zero there in every solver vector).  ``gm``: a precomputed interaction
    matrix (e.g. dense_green_matrix_pbc) instead of the isolated one."""
    nat = nx * ny * nz
    A = -(dense_green_matrix(nx, ny, nz, akd, dx, idipint) if gm is None else gm)
    adia = np.asarray(adia).reshape(3, nat)
    aoff = np.asarray(aoff).reshape(3, nat)
    idx = np.arange(nat)
    for c in range(3):
        A[c * nat + idx, c * nat + idx] += adia[c]
    a23, a31, a12 = aoff
    for (r, s, a) in ((0, 2, a31), (0, 1, a12), (1, 0, a12), (1, 2, a23), (2, 1, a23), (2, 0, a31)):
        A[r * nat + idx, s * nat + idx] += a
    if iocc is not None:
        occ3 = np.tile(np.asarray(iocc).reshape(-1) != 0, 3)
        A[~occ3, :] = 0
        A[:, ~occ3] = 0
    return A


# ----------------------------------------------------------------------------
# compact storage  (cu:120-134; scripts/check_compact_mapping.py)
# ----------------------------------------------------------------------------
def occupied_index(iocc) -> np.ndarray:
    """Occupied-site list in increasing j (cu:685-691)."""
    return np.flatnonzero(np.asarray(iocc).reshape(-1) != 0).astype(np.int64)


def gather_compact(full, occ, nat):
    """compact[c*NAT0+q] = full[c*NAT+occ[q]]  (cu:123-128)."""
    f = np.asarray(full).reshape(3, nat)
    return f[:, occ].reshape(-1)


def scatter_compact(compact, occ, nat):
    """full zero except full[c*NAT+occ[q]] = compact[c*NAT0+q]  (cu:129-134, cu:519)."""
    nat0 = occ.size
    c = np.asarray(compact).reshape(3, nat0)
    f = np.zeros((3, nat), dtype=c.dtype)
    f[:, occ] = c
    return f.reshape(-1)


# ----------------------------------------------------------------------------
# EVALQ  (cu:770-783, cu:919-955)
# ----------------------------------------------------------------------------
def evalq(e_c, p_c, adia_c, aoff_c, ak, e02):
    """Cabs, Cext, Cpha from compact [3][NAT0] arrays (cu:939-950)."""
    nat0 = np.asarray(p_c).size // 3
    p = np.asarray(p_c, np.complex128).reshape(3, nat0)
    e = np.asarray(e_c, np.complex128).reshape(3, nat0)
    ad = np.asarray(adia_c, np.complex128).reshape(3, nat0)
    ao = np.asarray(aoff_c, np.complex128).reshape(3, nat0)
    akk = float(np.sqrt(np.sum(np.asarray(ak, np.float64) ** 2)))
    rabs = akk ** 3 / 1.5
    ri = 1j * np.float64(np.float32(rabs))
    a23, a31, a12 = ao
    q = np.empty_like(p)
    q[0] = (ad[0] + ri) * p[0] + a31 * p[2] + a12 * p[1]
    q[1] = (ad[1] + ri) * p[1] + a12 * p[0] + a23 * p[2]
    q[2] = (ad[2] + ri) * p[2] + a23 * p[1] + a31 * p[0]
    zabs = np.vdot(p.reshape(-1), q.reshape(-1))
    zep = np.vdot(e.reshape(-1), p.reshape(-1))
    cabs = -4.0 * PI * akk * zabs.imag / e02
    cext = 4.0 * PI * akk * zep.imag / e02
    cpha = 2.0 * PI * akk * zep.real / e02
    return cabs, cext, cpha


# ----------------------------------------------------------------------------
# SCAT  (cu:849-875 directions, cu:793-845 sums, cu:957-1002 assembly)
# ----------------------------------------------------------------------------
def scat_integration_dirs(ak, e01, etasca, nat0):
    """DDSCAT angular quadrature for Csca/<cos>: returns (dirs[N,3], w[N,5]).
    theta grid and weights follow cu:849-875: x = k a_eff, theta0 = 2pi/(1+x),
    quadratic theta map, n_phi = round(4 pi sin(theta)/dtheta) >= 3, weight
    dOmega/n_phi, phi offset alternating with the theta-ring parity; basis
    a1 || Re(E01), a2 = k_hat x a1."""
    ak = np.asarray(ak, np.float64)
    akk = float(np.sqrt(np.sum(ak * ak)))
    a1 = np.array([complex(v).real for v in e01], dtype=np.float64)
    einc = float(np.sqrt(np.sum(a1 * a1)))
    if not einc > 0:
        raise ValueError("CXE01 is pure imaginary")
    a1 = a1 * akk / einc
    a2 = np.array([ak[1] * a1[2] - ak[2] * a1[1], ak[2] * a1[0] - ak[0] * a1[2],
                   ak[0] * a1[1] - ak[1] * a1[0]]) / akk
    x = akk * (3.0 * nat0 / (4.0 * PI)) ** (1.0 / 3.0)
    theta0 = 2.0 * PI / (1.0 + x)
    afac = 1.0
    ntheta = 1 + int(_round_half_away((2.0 * (3.0 + x) / etasca) * (1.0 + afac / (PI + theta0))
                                      / (1.0 + afac * theta0 / ((PI + theta0) ** 2))))
    theta = 0.0
    thetau = 0.0
    dirs, wts = [], []
    for ic in range(1, ntheta + 1):
        thetal = theta
        theta = thetau
        if ic < ntheta:
            si = ic * (PI + afac * PI / (theta0 + PI)) / (ntheta - 1)
            term = si - afac - theta0
            thetau = 0.5 * (term + math.sqrt(term * term + 4.0 * theta0 * si))
        else:
            thetau = PI
        domega = 2.0 * PI * (math.cos(0.5 * (thetal + theta)) - math.cos(0.5 * (theta + thetau)))
        ct, st = math.cos(theta), math.sin(theta)
        if ic == 1 or ic == ntheta:
            nphi = 1
        else:
            nphi = int(_round_half_away(4.0 * PI * st / (thetau - thetal)))
            nphi = max(nphi, 3)
        w = domega / nphi
        dphi = 2.0 * PI / nphi
        for ip in range(1, nphi + 1):
            phi = dphi * (ip - 0.5 * (ic % 2))
            sp, cp = math.sin(phi), math.cos(phi)
            k = ct * ak + st * (cp * a1 + sp * a2)
            dirs.append(k.astype(np.float32).astype(np.float64))
            wts.append((w, w * ct, w * st * cp, w * st * sp, w * ct * ct))
    return np.array(dirs), np.array(wts)


def _round_half_away(v):
    return math.floor(v + 0.5) if v >= 0 else -math.floor(-v + 0.5)


def site_positions(occ, nx, ny, dx, x0):
    """r_j = ((ix+1)+x0_1)dx_1, ... for compact site list (cu:800-802)."""
    occ = np.asarray(occ)
    ix = occ % nx
    iy = (occ // nx) % ny
    iz = occ // (nx * ny)
    r = np.stack([(ix + 1 + x0[0]) * dx[0], (iy + 1 + x0[1]) * dx[1], (iz + 1 + x0[2]) * dx[2]], axis=1)
    return r.astype(np.float64)


def farfield_sums(p_c, occ, nx, ny, dx, x0, dirs):
    """S[dir] = sum_j P_j exp(-i k_s . r_j)  (cu:793-822), FP64."""
    nat0 = occ.size
    p = np.asarray(p_c, np.complex128).reshape(3, nat0)
    r = site_positions(occ, nx, ny, dx, x0)
    out = np.empty((len(dirs), 3), dtype=np.complex128)
    for i0 in range(0, len(dirs), 64):
        d = np.asarray(dirs[i0:i0 + 64], np.float64)
        ph = np.exp(-1j * (d @ r.T))            # [nd, nat0]
        out[i0:i0 + 64] = ph @ p.T
    return out


def scat(p_c, occ, nx, ny, ak, dx, x0, e01, etasca, e02, aks=None, em1=None, em2=None):
    """Csca, Csca*g(3), Csca*<cos^2>, Cbk and selected F1/F2 (cu:957-1002)."""
    ak = np.asarray(ak, np.float64)
    ak2 = float(np.sum(ak * ak))
    akk = math.sqrt(ak2)
    dirs, w = scat_integration_dirs(ak, e01, etasca, occ.size)
    S = farfield_sums(p_c, occ, nx, ny, dx, x0, dirs)
    n = dirs / akk
    ndotp = np.sum(n * S, axis=1)
    E = S - n * ndotp[:, None]
    es2 = np.sum(np.abs(E) ** 2, axis=1)
    acc = w.T @ es2
    fac = ak2 * ak2 / e02
    res = dict(csca=fac * acc[0], cscag=fac * acc[1:4], cscag2=fac * acc[4], cbksca=fac * es2[-1],
               navg=len(dirs))
    if aks is not None and len(aks):
        Ss = farfield_sums(p_c, occ, nx, ny, dx, x0, np.asarray(aks, np.float64))
        term = akk ** 3 / math.sqrt(e02)
        res["f1"] = term * np.sum(np.asarray(em1, np.float64) * Ss, axis=1)
        res["f2"] = term * np.sum(np.asarray(em2, np.float64) * Ss, axis=1)
    return res
