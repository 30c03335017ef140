"""GPU parity of the Green-tensor setup and the FFT3D MATVEC against the FP64
oracle (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:104-118, 644-652).

Tolerances: Green octant within 2e-6 of its largest entry (FP64 build, one
rounding to complex64); MATVEC relative L2 error <= 1e-5 (the stated
MATVEC parity bar) against the FP64 oracle on the same complex64 inputs."""
import numpy as np
import pytest

from oracle import ddscat_oracle as O
from ddscat_hip import targets as T

pytestmark = pytest.mark.gpu

MATVEC_TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.complex128) - b) / np.linalg.norm(b))


def random_case(rng, nx, ny, nz, fill=0.7):
    nat = nx * ny * nz
    iocc = (rng.random(nat) < fill).astype(np.int16)
    iocc[0] = 1
    mask = np.tile(iocc != 0, 3)
    cplx = lambda: (rng.normal(size=3 * nat) + 1j * rng.normal(size=3 * nat)).astype(np.complex64)
    x = cplx() * mask
    adia = cplx()
    aoff = cplx() * 0.3
    return iocc, x, adia, aoff


@pytest.mark.parametrize("dims,akd,dx", [
    ((4, 4, 4), 0.3, (1.0, 1.0, 1.0)),
    ((5, 6, 4), 0.41, (1.0, 1.0, 1.0)),
    ((8, 6, 10), 0.2, (1.1, 0.9, 1.0 / 0.99)),
    ((16, 12, 20), 0.15, (1.0, 1.0, 1.0)),
])
def test_green_octant_matches_oracle(lib, dims, akd, dx):
    nx, ny, nz = dims
    rng = np.random.default_rng(7)
    iocc, *_ = random_case(rng, nx, ny, nz)
    ak = (akd, 0.0, 0.0)
    lib.prepare_green(nx, ny, nz, iocc, ak, dx)
    g = lib.get_green().reshape(6, nz + 1, ny + 1, nx + 1)
    ref = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)), tuple(np.float64(np.float32(v)) for v in dx))
    err = np.abs(g - ref).max() / np.abs(ref).max()
    print(f"green octant {dims}: max err / max = {err:.3e}")
    assert err < 2e-6


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 5, 4), (8, 8, 8), (10, 6, 9), (12, 15, 16), (24, 20, 18)])
@pytest.mark.parametrize("cwhat", ["N", "C"])
def test_matvec_matches_oracle(lib, dims, cwhat):
    nx, ny, nz = dims
    rng = np.random.default_rng(sum(dims) + (cwhat == "C"))
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    akd = 0.37
    lib.prepare_green(nx, ny, nz, iocc, (0.0, akd, 0.0))
    y = lib.apply(cwhat, x, adia, aoff)
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)))
    yref = O.matvec_fft(cwhat, x.astype(np.complex128), adia.astype(np.complex128), aoff.astype(np.complex128),
                        goct, nx, ny, nz, iocc)
    err = rel_l2(y, yref)
    print(f"matvec {dims} {cwhat}: rel L2 = {err:.3e}")
    assert err < MATVEC_TOL
    assert np.all(y.reshape(3, -1)[:, iocc == 0] == 0)


@pytest.mark.parametrize("dims", [(4, 6, 384), (6, 384, 4), (384, 4, 6), (5, 4, 320), (4, 3, 512)])
def test_wide_lds_tiles_match_oracle(lib, dims):
    """Padded axes of 640-1024 points (no register shape): the LDS-staged y
    and z passes switch to 1024-thread wide tiles (8-16 columns per tile)."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * ny * nz)
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    akd = 0.23
    lib.prepare_green(nx, ny, nz, iocc, (akd, 0.0, 0.0))
    info = lib.info()
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)))
    for cwhat in ("N", "C"):
        y = lib.apply(cwhat, x, adia, aoff)
        yref = O.matvec_fft(cwhat, x.astype(np.complex128), adia.astype(np.complex128),
                            aoff.astype(np.complex128), goct, nx, ny, nz, iocc)
        err = rel_l2(y, yref)
        print(f"wide tiles {dims} {cwhat} (cy={info['cy']}, cz={info['cz']}): rel L2 = {err:.3e}")
        assert err < MATVEC_TOL
    if 2 * ny >= 640:
        assert info["cy"] >= 8
    if 2 * nz >= 640:
        assert info["cz"] >= 4


def test_matvec_dense_bruteforce(lib):
    """GPU MATVEC equals the dense 3NAT x 3NAT operator built from the closed form."""
    nx, ny, nz = 4, 3, 5
    rng = np.random.default_rng(11)
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz, fill=0.6)
    akd = 0.52
    lib.prepare_green(nx, ny, nz, iocc, (0.0, 0.0, akd))
    y = lib.apply("N", x, adia, aoff)
    A = O.dense_operator(nx, ny, nz, np.float64(np.float32(akd)), adia.astype(np.complex128),
                         aoff.astype(np.complex128), iocc)
    assert rel_l2(y, A @ x.astype(np.complex128)) < MATVEC_TOL


def test_matvec_sphere_d32(lib):
    """The small workload: sphere of diameter 32 (17,256 dipoles), m=1.33+0.01i, x=2."""
    tg = T.sphere(32)
    kd = T.kd_for_size_parameter(tg, 2.0)
    khat = np.array([1.0, 0.0, 0.0])
    adia, aoff = T.operators(tg, 1.33 + 0.01j, kd, khat, np.array([0, 1, 0]))
    rng = np.random.default_rng(3)
    x = ((rng.normal(size=3 * tg.nat) + 1j * rng.normal(size=3 * tg.nat)) * np.tile(tg.iocc != 0, 3)).astype(np.complex64)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    y = lib.apply("N", x, adia, aoff)
    goct = O.green_octant(tg.nx, tg.ny, tg.nz, np.float64(np.float32(kd)))
    yref = O.matvec_fft("N", x, adia, aoff, goct, tg.nx, tg.ny, tg.nz, tg.iocc)
    err = rel_l2(y, yref)
    print(f"sphere D=32 matvec rel L2 = {err:.3e}")
    assert err < MATVEC_TOL


def test_prepare_with_host_green(lib):
    """prepare_f32 path: the CPU-built CXZC octant uploaded as is."""
    nx, ny, nz = 6, 5, 7
    rng = np.random.default_rng(5)
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    goct = O.green_octant(nx, ny, nz, 0.25)
    lib.prepare(nx, ny, nz, goct.astype(np.complex64), iocc)
    y = lib.apply("N", x, adia, aoff)
    yref = O.matvec_fft("N", x, adia, aoff, goct.astype(np.complex64).astype(np.complex128), nx, ny, nz, iocc)
    assert rel_l2(y, yref) < MATVEC_TOL


def test_compact_device_layout_matches_full(lib):
    """ddscat_hip_apply_device on compact [3][NAT0] vectors equals the full-layout MATVEC."""
    import torch
    nx, ny, nz = 12, 10, 14
    rng = np.random.default_rng(9)
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz, fill=0.55)
    lib.prepare_green(nx, ny, nz, iocc, (0.3, 0.0, 0.0))
    yfull = lib.apply("N", x, adia, aoff)
    occ = np.flatnonzero(iocc)
    comp = lambda v: np.ascontiguousarray(v.reshape(3, -1)[:, occ].reshape(-1))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()
    xc, ac, oc = dev(comp(x)), dev(comp(adia)), dev(comp(aoff))
    yc = torch.empty_like(xc)
    torch.cuda.synchronize()
    lib.apply_device("N", xc.data_ptr(), yc.data_ptr(), ac.data_ptr(), oc.data_ptr(), layout=1)
    lib.sync()
    got = yc.cpu().numpy().view(np.complex64)
    assert rel_l2(got, comp(yfull).astype(np.complex128)) < 1e-6


def test_apply_before_prepare_and_bad_args(lib):
    lib.release()
    rc = lib.lib.ddscat_cuda_apply_f32(ord("N"), None, None, None, None)
    assert rc == 10
    iocc = np.ones(8, np.int16)
    lib.prepare_green(2, 2, 2, iocc, (0.1, 0, 0))
    x = np.zeros(24, np.complex64)
    rc = lib.lib.ddscat_cuda_apply_f32(ord("X"), x.ctypes.data, x.ctypes.data, x.ctypes.data, x.ctypes.data)
    assert rc == 11


@pytest.mark.parametrize("dims", [(32, 64, 128), (32, 20, 64), (64, 8, 32), (128, 4, 6)])
@pytest.mark.parametrize("cwhat", ["N", "C"])
def test_register_fft_paths_match_oracle(lib, dims, cwhat):
    """Axes of padded length 64/128/256 run the register-resident FFT kernels."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + 7 * ny + 13 * nz + (cwhat == "C"))
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    akd = 0.21
    lib.prepare_green(nx, ny, nz, iocc, (0.0, 0.0, akd))
    y = lib.apply(cwhat, x, adia, aoff)
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)))
    yref = O.matvec_fft(cwhat, x, adia, aoff, goct, nx, ny, nz, iocc)
    err = rel_l2(y, yref)
    print(f"register path {dims} {cwhat}: rel L2 = {err:.3e}")
    assert err < MATVEC_TOL


@pytest.mark.parametrize("target", ["sphere", "random"])
def test_register_and_generic_kernels_agree_on_compact_vectors(lib, target):
    """Both FFT kernel families, compact device layout; sphere rows are
    intervals (index-free addressing), random rows use the site map."""
    import torch
    if target == "sphere":
        tg = T.sphere(32)
        iocc = tg.iocc
        nx = ny = nz = 32
    else:
        nx, ny, nz = 32, 16, 64
        iocc = (np.random.default_rng(4).random(nx * ny * nz) < 0.6).astype(np.int16)
    occ = np.flatnonzero(iocc)
    rng = np.random.default_rng(12)
    n0 = 3 * occ.size
    cx = lambda s=1.0: (s * (rng.normal(size=n0) + 1j * rng.normal(size=n0))).astype(np.complex64)
    x, ad, ao = cx(), cx(), cx(0.2)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()
    outs = []
    for generic in (0, 1):
        lib.set_fft_kernels(generic)
        lib.prepare_green(nx, ny, nz, iocc, (0.3, 0.1, 0.0))
        xd, ad_d, ao_d = dev(x), dev(ad), dev(ao)
        yd = torch.empty_like(xd)
        torch.cuda.synchronize()
        lib.apply_device("N", xd.data_ptr(), yd.data_ptr(), ad_d.data_ptr(), ao_d.data_ptr(), layout=1)
        lib.sync()
        outs.append(yd.cpu().numpy().view(np.complex64).astype(np.complex128))
    lib.set_fft_kernels(0)
    xf = np.zeros(3 * nx * ny * nz, complex)
    xf.reshape(3, -1)[:, occ] = x.reshape(3, -1)
    adf = np.zeros_like(xf); adf.reshape(3, -1)[:, occ] = ad.reshape(3, -1)
    aof = np.zeros_like(xf); aof.reshape(3, -1)[:, occ] = ao.reshape(3, -1)
    kd = float(np.linalg.norm(np.float32([0.3, 0.1, 0.0])))
    yref = O.matvec_fft("N", xf, adf, aof, O.green_octant(nx, ny, nz, kd), nx, ny, nz, iocc)
    yref = yref.reshape(3, -1)[:, occ].reshape(-1)
    for y in outs:
        assert rel_l2(y, yref) < MATVEC_TOL
    assert rel_l2(outs[0], outs[1]) < 1e-6


_ROT = [(16, 24, 30), (36, 40, 48), (50, 54, 6), (60, 72, 4), (75, 80, 2), (90, 96, 2), (120, 3, 2)]
GENERAL_SHAPE_DIMS = [d[k:] + d[:k] for d in _ROT for k in range(3)]


@pytest.mark.parametrize("dims", GENERAL_SHAPE_DIMS)
def test_general_register_shapes_match_oracle(lib, dims):
    """Padded axes 32 ... 240 with factors 2, 3, 5 (register shapes with
    R1 not dividing R2 and idle second-step lanes), each length on every
    axis so passes 1-5 all run it; full layout ('N') and compact layout ('C')."""
    import torch
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 10007 + ny * 101 + nz)
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    akd = 0.23
    lib.prepare_green(nx, ny, nz, iocc, (0.0, akd, 0.0))
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)))
    y = lib.apply("N", x, adia, aoff)
    yref = O.matvec_fft("N", x, adia, aoff, goct, nx, ny, nz, iocc)
    err_full = rel_l2(y, yref)
    occ = np.flatnonzero(iocc)
    comp = lambda v: np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, occ].reshape(-1))
    dev = lambda a: torch.from_numpy(comp(a).view(np.float32)).cuda()
    xd, ad, ao = dev(x), dev(adia), dev(aoff)
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    lib.apply_device("C", xd.data_ptr(), yd.data_ptr(), ad.data_ptr(), ao.data_ptr(), layout=1)
    lib.sync()
    yc = yd.cpu().numpy().view(np.complex64)
    ycref = comp(O.matvec_fft("C", x, adia, aoff, goct, nx, ny, nz, iocc))
    err_c = rel_l2(yc, ycref)
    print(f"{dims}: full N {err_full:.2e}, compact C {err_c:.2e}")
    assert err_full < MATVEC_TOL and err_c < MATVEC_TOL


@pytest.mark.parametrize("backend", ["fft3d", "slices"])
def test_nearfield_matches_oracle_on_all_sites(lib, backend):
    """E_inc + E_sca on every lattice site, empty sites included, equals the
    oracle's convolution of P with the Green tensor (FFT MATVEC with the local
    terms off and no occupancy mask, sign-flipped), on either backend."""
    import torch
    from ddscat_hip import BACKEND_FFT3D, BACKEND_SLICES
    nx, ny, nz = 12, 10, 14
    rng = np.random.default_rng(33)
    iocc, x, _, _ = random_case(rng, nx, ny, nz, fill=0.5)
    akd = 0.31
    lib.release()
    lib.set_backend(BACKEND_SLICES if backend == "slices" else BACKEND_FFT3D)
    try:
        _nearfield_case(lib, nx, ny, nz, iocc, x, akd, rng)
    finally:
        lib.release()
        lib.set_backend(BACKEND_FFT3D)


def _nearfield_case(lib, nx, ny, nz, iocc, x, akd, rng):
    import torch
    lib.prepare_green(nx, ny, nz, iocc, (akd, 0.0, 0.0))
    occ = np.flatnonzero(iocc)
    p_c = np.ascontiguousarray(x.reshape(3, -1)[:, occ].reshape(-1))
    einc = (rng.normal(size=x.size) + 1j * rng.normal(size=x.size)).astype(np.complex64)
    pd = torch.from_numpy(p_c).cuda()
    ed = torch.from_numpy(einc).cuda()
    out = torch.empty_like(ed)
    torch.cuda.synchronize()
    lib.nearfield(pd.data_ptr(), out.data_ptr(), ed.data_ptr())
    got = out.cpu().numpy()
    zero = np.zeros_like(x, dtype=complex)
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)))
    esca = -O.matvec_fft("N", x.astype(complex), zero, zero, goct, nx, ny, nz, None)
    ref = einc + esca
    err = rel_l2(got, ref)
    void = np.tile(iocc == 0, 3)
    err_void = rel_l2(got[void], ref[void])
    print(f"near field rel L2 {err:.2e} (empty sites {err_void:.2e})")
    assert err < MATVEC_TOL and err_void < MATVEC_TOL
    lib.nearfield(pd.data_ptr(), out.data_ptr())
    assert rel_l2(out.cpu().numpy(), esca) < MATVEC_TOL


@pytest.mark.parametrize("dims", [(128, 8, 6), (6, 128, 4), (4, 6, 128), (32, 32, 32), (72, 6, 72), (20, 30, 40),
                                  (64, 6, 10), (384, 6, 4), (6, 384, 4), (4, 6, 384), (256, 4, 6), (6, 256, 4), (4, 6, 256),
                                  (100, 6, 4), (4, 100, 6), (6, 4, 100), (108, 4, 6), (6, 108, 4), (4, 6, 108),
                                  (6, 4, 48), (4, 6, 60), (6, 4, 75), (4, 6, 80), (6, 4, 90), (4, 6, 96), (6, 4, 120)])
def test_every_kernel_candidate_matches_oracle(lib, dims):
    """Each FFT kernel the autotuner may pick (every register shape of the
    pass's axis length, the lane-pair pass 3, and the LDS-staged kernel),
    forced pass by pass, gives the oracle MATVEC."""
    import torch
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    iocc, x, adia, aoff = random_case(rng, nx, ny, nz)
    akd = 0.29
    lib.prepare_green(nx, ny, nz, iocc, (akd, 0.0, 0.0))
    goct = O.green_octant(nx, ny, nz, np.float64(np.float32(akd)))
    yref = O.matvec_fft("N", x, adia, aoff, goct, nx, ny, nz, iocc)
    tuned = lib.info()
    base = [tuned[f"kernel_pass{p}"] for p in range(1, 6)]
    checked = 0
    for p in range(1, 6):
        kind = -1
        while lib.set_pass_kernel(p, kind, check=False) == 0:
            y = lib.apply("N", x, adia, aoff)
            err = rel_l2(y, yref)
            assert err < MATVEC_TOL, (dims, p, kind, err)
            checked += 1
            kind += 1
        lib.set_pass_kernel(p, base[p - 1])
    print(f"{dims}: {checked} kernel choices checked, tuned {base}")
    assert checked >= 6  # at least one register choice besides the LDS kernels


@pytest.mark.parametrize("dims", [(384, 5, 4), (256, 4, 5)])
@pytest.mark.parametrize("rows", ["interval", "sitemap"])
@pytest.mark.parametrize("cwhat", ["N", "C"])
def test_pass5_kernels_on_compact_vectors(lib, dims, rows, cwhat):
    """Every pass-5 kernel of a 768- or 512-point x axis (the 3 x 256 and
    2 x 256 decimated register kernels and the LDS-staged one) on compact
    [3][NAT0] device vectors: occupied sites forming one interval per row
    (index-free addressing) or scattered (site map), 'N' and 'C'."""
    import torch
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + 5 * ny + 11 * nz + (rows == "sitemap"))
    if rows == "interval":
        lo = rng.integers(0, nx // 2, size=ny * nz)
        ln = rng.integers(0, nx - lo)
        lo[0], ln[0] = 0, max(int(ln[0]), 1)  # row 0 starts at site 0, as NAT0 > 0 needs a site
        ix = np.arange(nx)
        iocc = ((ix >= lo[:, None]) & (ix < (lo + ln)[:, None])).astype(np.int16).reshape(-1)
    else:
        iocc = (rng.random(nx * ny * nz) < 0.6).astype(np.int16)
        iocc[0] = 1
    occ = np.flatnonzero(iocc)
    n0 = 3 * occ.size
    cx = lambda s=1.0: (s * (rng.normal(size=n0) + 1j * rng.normal(size=n0))).astype(np.complex64)
    x, ad, ao = cx(), cx(), cx(0.3)
    akd = (0.27, 0.05, 0.1)
    lib.prepare_green(nx, ny, nz, iocc, akd)
    full = lambda v: np.zeros(3 * nx * ny * nz, complex)
    xf, adf, aof = full(0), full(0), full(0)
    xf.reshape(3, -1)[:, occ] = x.reshape(3, -1)
    adf.reshape(3, -1)[:, occ] = ad.reshape(3, -1)
    aof.reshape(3, -1)[:, occ] = ao.reshape(3, -1)
    kd = float(np.linalg.norm(np.float32(akd)))
    yref = O.matvec_fft(cwhat, xf, adf, aof, O.green_octant(nx, ny, nz, kd), nx, ny, nz, iocc)
    yref = yref.reshape(3, -1)[:, occ].reshape(-1)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()
    xd, add, aod = dev(x), dev(ad), dev(ao)
    yd = torch.empty_like(xd)
    base = lib.info()["kernel_pass5"]
    kind, checked = -1, 0
    while lib.set_pass_kernel(5, kind, check=False) == 0:
        torch.cuda.synchronize()
        lib.apply_device(cwhat, xd.data_ptr(), yd.data_ptr(), add.data_ptr(), aod.data_ptr(), layout=1)
        lib.sync()
        err = rel_l2(yd.cpu().numpy().view(np.complex64), yref)
        assert err < MATVEC_TOL, (dims, rows, cwhat, kind, err)
        checked += 1
        kind += 1
    lib.set_pass_kernel(5, base)
    assert checked >= 3  # the LDS-staged kernel and two decimated tiles
