"""ctypes binding of the drop-in C ABI (include/ddscat_matvec_hip.h).

Mirrors what the reference's Fortran bridge does through ISO_C_BINDING
(ddscat_cuda_unified/fortran/cuda_matvec_bridge.f90:14-115): host arrays in,
host arrays out, integer return codes with the library's last_error text.
The binding refuses to run if the HIP library is not built: there is no CPU
fallback on this path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_DIR = PKG_ROOT / "lib"
LIB_FFT3D = LIB_DIR / "libddscat_matvec_cuda.so"
LIB_SLICES = LIB_DIR / "libddscat_matvec_cuda_slice.so"

SOLVERS = ("gpbicg", "qmrccg", "pbcgst", "pbcgs2", "petrkp", "bicgstab4", "gpbicgstab2", "gpbicgstab4")
REFERENCE_SYMBOLS = (
    ["ddscat_cuda_prepare_green_f32", "ddscat_cuda_prepare_f32", "ddscat_cuda_apply_f32"]
    + [f"ddscat_cuda_solve_{s}_f32" for s in SOLVERS]
    + ["ddscat_cuda_evalq_f32", "ddscat_cuda_scat_f32", "ddscat_cuda_release", "ddscat_cuda_last_error"]
)
EXTENSION_SYMBOLS = [
    "ddscat_hip_set_backend", "ddscat_hip_set_precision", "ddscat_hip_apply_device",
    "ddscat_hip_get_green", "ddscat_hip_info", "ddscat_hip_stream", "ddscat_hip_sync",
    "ddscat_hip_time_matvec", "ddscat_hip_set_fft_kernels", "ddscat_hip_solve_device",
    "ddscat_hip_solve_batch", "ddscat_hip_apply_batch", "ddscat_hip_evalq_batch",
    "ddscat_hip_cabs_adjoint", "ddscat_hip_setup_rhs", "ddscat_hip_set_autotune",
    "ddscat_hip_nearfield", "ddscat_hip_set_pass_kernel", "ddscat_hip_scat_matrix", "ddscat_hip_force_restart",
    "ddscat_hip_solve_log", "ddscat_hip_set_polarizability", "ddscat_hip_scat_f32_noalloc",
    "ddscat_hip_torque", "ddscat_hip_set_reliable_mode", "ddscat_hip_select_target", "ddscat_hip_solve_targets",
    "ddscat_hip_poison_workspace",
]
# method ids of ddscat_hip_solve_device
SOLVER_IDS = {"pbcgs2": 0, "qmrccg": 1, "petrkp": 2, "gpbicg": 3, "pbcgst": 4, "bicgstab4": 5,
              "gpbicgstab2": 6, "gpbicgstab4": 7}
INFO_FIELDS = ("nx", "ny", "nz", "nat", "nat0", "green_count", "matvec_count", "backend", "precision",
               "gx", "gy", "gz", "rx_fwd", "cy", "cz", "rx_inv", "green_build_us",
               "kernel_pass1", "kernel_pass2", "kernel_pass3", "kernel_pass4", "kernel_pass5", "autotuned",
               "restarts")

BACKEND_FFT3D, BACKEND_SLICES = 0, 1
PREC_MIXED, PREC_FP32 = 0, 1


class DDSCATError(RuntimeError):
    def __init__(self, where: str, code: int, text: str):
        super().__init__(f"{where} failed with code {code}: {text}")
        self.code = code
        self.text = text


def _c64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.complex64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} complex values, got {a.size}")
    return a


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Library:
    """One loaded copy of the plugin.  The C state is a process-wide singleton,
    exactly as in the reference (one prepared target at a time)."""

    def __init__(self, path: str | os.PathLike | None = None, slices: bool = False):
        if path is None:
            path = os.environ.get("DDSCAT_CUDA_DLL") or (LIB_SLICES if slices else LIB_FFT3D)
        path = Path(path)
        if not path.exists():
            raise FileNotFoundError(
                f"HIP library {path} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        self.path = path
        # torch ships its own HIP runtime.  Loading it first lets this library
        # bind to that runtime; loaded the other way round, the process ends up
        # with a runtime that finds no device after torch initialises the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        self.lib = C.CDLL(str(path), mode=C.RTLD_LOCAL)
        L = self.lib
        i, f, vp, cp = C.c_int, C.c_float, C.c_void_p, C.c_char_p
        L.ddscat_cuda_prepare_green_f32.argtypes = [i] * 7 + [f] * 9 + [vp]
        L.ddscat_cuda_prepare_green_f32.restype = i
        L.ddscat_cuda_prepare_f32.argtypes = [i] * 6 + [vp, vp]
        L.ddscat_cuda_prepare_f32.restype = i
        L.ddscat_cuda_apply_f32.argtypes = [i, vp, vp, vp, vp]
        L.ddscat_cuda_apply_f32.restype = i
        for s in SOLVERS:
            fn = getattr(L, f"ddscat_cuda_solve_{s}_f32")
            fn.argtypes = [vp, vp, vp, vp, i, i, f, vp, vp, vp]
            fn.restype = i
        L.ddscat_cuda_evalq_f32.argtypes = [i, vp, vp, vp, vp, f, f, f, f, vp, vp, vp]
        L.ddscat_cuda_evalq_f32.restype = i
        L.ddscat_cuda_scat_f32.argtypes = [i, i, vp, vp, vp, vp, vp, f, f, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.ddscat_cuda_scat_f32.restype = i
        L.ddscat_hip_scat_f32_noalloc.argtypes = list(L.ddscat_cuda_scat_f32.argtypes)
        L.ddscat_hip_scat_f32_noalloc.restype = i
        L.ddscat_hip_torque.argtypes = [vp, vp, f, f, vp, vp, vp, vp, vp]
        L.ddscat_hip_torque.restype = i
        L.ddscat_hip_set_reliable_mode.argtypes = [i]
        L.ddscat_hip_set_reliable_mode.restype = i
        L.ddscat_hip_select_target.argtypes = [i]
        L.ddscat_hip_select_target.restype = i
        L.ddscat_hip_solve_targets.argtypes = [i, i, vp, vp, vp, vp, vp, vp, i, i, f, vp, vp, vp, i]
        L.ddscat_hip_solve_targets.restype = i
        L.ddscat_cuda_release.argtypes = []
        L.ddscat_cuda_release.restype = i
        L.ddscat_cuda_last_error.argtypes = []
        L.ddscat_cuda_last_error.restype = cp
        L.ddscat_hip_set_backend.argtypes = [i]
        L.ddscat_hip_set_backend.restype = i
        L.ddscat_hip_set_precision.argtypes = [i]
        L.ddscat_hip_set_precision.restype = i
        L.ddscat_hip_apply_device.argtypes = [i, vp, vp, vp, vp, i, vp]
        L.ddscat_hip_apply_device.restype = i
        L.ddscat_hip_get_green.argtypes = [vp]
        L.ddscat_hip_get_green.restype = i
        L.ddscat_hip_info.argtypes = [vp, i]
        L.ddscat_hip_info.restype = i
        L.ddscat_hip_stream.argtypes = []
        L.ddscat_hip_stream.restype = vp
        L.ddscat_hip_sync.argtypes = []
        L.ddscat_hip_sync.restype = i
        L.ddscat_hip_time_matvec.argtypes = [i, vp, vp, vp, vp, i, i, vp]
        L.ddscat_hip_time_matvec.restype = i
        L.ddscat_hip_set_fft_kernels.argtypes = [i]
        L.ddscat_hip_set_fft_kernels.restype = i
        L.ddscat_hip_solve_device.argtypes = [i, vp, vp, vp, vp, i, i, f, vp, vp, vp]
        L.ddscat_hip_solve_device.restype = i
        L.ddscat_hip_solve_batch.argtypes = [i, i, vp, vp, vp, vp, i, i, f, vp, vp, vp, i]
        L.ddscat_hip_solve_batch.restype = i
        L.ddscat_hip_apply_batch.argtypes = [i, i, vp, vp, vp, vp, i, vp]
        L.ddscat_hip_apply_batch.restype = i
        L.ddscat_hip_evalq_batch.argtypes = [i, vp, vp, vp, vp, i, vp, f, vp]
        L.ddscat_hip_evalq_batch.restype = i
        L.ddscat_hip_cabs_adjoint.argtypes = [i, i, vp, vp, vp, i, vp, f, i, f, vp, vp, vp, vp]
        L.ddscat_hip_cabs_adjoint.restype = i
        L.ddscat_hip_setup_rhs.argtypes = [i, vp, vp, vp, vp, vp, i, vp, vp, vp]
        L.ddscat_hip_setup_rhs.restype = i
        L.ddscat_hip_set_autotune.argtypes = [i]
        L.ddscat_hip_set_autotune.restype = i
        L.ddscat_hip_nearfield.argtypes = [vp, vp, vp]
        L.ddscat_hip_nearfield.restype = i
        L.ddscat_hip_scat_matrix.argtypes = [i, vp, i, vp, vp, vp, vp, vp, vp, vp, f, vp, vp]
        L.ddscat_hip_scat_matrix.restype = i
        L.ddscat_hip_force_restart.argtypes = []
        L.ddscat_hip_force_restart.restype = i
        L.ddscat_hip_poison_workspace.argtypes = []
        L.ddscat_hip_poison_workspace.restype = i
        L.ddscat_hip_solve_log.argtypes = [i, vp, vp, i, vp, i, vp]
        L.ddscat_hip_solve_log.restype = i
        L.ddscat_hip_set_polarizability.argtypes = [i]
        L.ddscat_hip_set_polarizability.restype = i
        L.ddscat_hip_set_pass_kernel.argtypes = [i, i]
        L.ddscat_hip_set_pass_kernel.restype = i
        self.nat = self.nat0 = 0

    # ---------------------------------------------------------------- errors
    def last_error(self) -> str:
        t = self.lib.ddscat_cuda_last_error()
        return t.decode() if t else ""

    def _check(self, where, rc):
        if rc != 0:
            raise DDSCATError(where, rc, self.last_error())
        return rc

    # --------------------------------------------------------------- prepare
    def prepare_green(self, nx, ny, nz, iocc, ak, dx=(1.0, 1.0, 1.0), ipbc=0, idipint=0, gamma=0.0,
                      pyd=0.0, pzd=0.0, nat=None, nat0=None, check=True):
        iocc = np.ascontiguousarray(iocc, dtype=np.int16)
        nat = nx * ny * nz if nat is None else nat
        nat0 = int(np.count_nonzero(iocc)) if nat0 is None else nat0
        rc = self.lib.ddscat_cuda_prepare_green_f32(nx, ny, nz, ipbc, nat, nat0, idipint, gamma, pyd, pzd,
                                                    float(ak[0]), float(ak[1]), float(ak[2]),
                                                    float(dx[0]), float(dx[1]), float(dx[2]), _p(iocc))
        if rc == 0:
            self.nat, self.nat0 = nat, nat0
        return self._check("prepare_green", rc) if check else rc

    def prepare(self, nx, ny, nz, green, iocc, ipbc=0, nat=None, nat0=None, check=True):
        iocc = np.ascontiguousarray(iocc, dtype=np.int16)
        nat = nx * ny * nz if nat is None else nat
        nat0 = int(np.count_nonzero(iocc)) if nat0 is None else nat0
        g = _c64(green)
        rc = self.lib.ddscat_cuda_prepare_f32(nx, ny, nz, ipbc, nat, nat0, _p(g), _p(iocc))
        if rc == 0:
            self.nat, self.nat0 = nat, nat0
        return self._check("prepare", rc) if check else rc

    def release(self):
        return self.lib.ddscat_cuda_release()

    # ----------------------------------------------------------------- apply
    def apply(self, cwhat, x, adia, aoff):
        n = 3 * self.nat
        x, adia, aoff = _c64(x, n), _c64(adia, n), _c64(aoff, n)
        y = np.empty(n, dtype=np.complex64)
        cw = ord(cwhat) if isinstance(cwhat, str) else int(cwhat)
        self._check("apply", self.lib.ddscat_cuda_apply_f32(cw, _p(x), _p(y), _p(adia), _p(aoff)))
        return y

    # ---------------------------------------------------------------- solve
    def solve(self, name, b, adia, aoff, x0=None, maxit=300, tol=1e-5, check=True):
        """Returns (x, itno, tole, ncompte, rc)."""
        n = 3 * self.nat
        b, adia, aoff = _c64(b, n), _c64(adia, n), _c64(aoff, n)
        x = np.zeros(n, dtype=np.complex64) if x0 is None else _c64(x0, n).copy()
        it, nc = C.c_int(0), C.c_int(0)
        te = C.c_float(0.0)
        fn = getattr(self.lib, f"ddscat_cuda_solve_{name}_f32")
        rc = fn(_p(b), _p(x), _p(adia), _p(aoff), n, int(maxit), float(tol), C.byref(it), C.byref(te), C.byref(nc))
        if check:
            self._check(f"solve_{name}", rc)
        return x, it.value, te.value, nc.value, rc

    def solve_device(self, name, b_ptr, x_ptr, adia_ptr, aoff_ptr, maxit=300, tol=1e-5, check=True):
        """Solver on compact [3][NAT0] device vectors (x in/out).  Returns (itno, tole, ncompte, rc)."""
        it, nc = C.c_int(0), C.c_int(0)
        te = C.c_float(0.0)
        rc = self.lib.ddscat_hip_solve_device(SOLVER_IDS[name], C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                              C.c_void_p(adia_ptr), C.c_void_p(aoff_ptr), 3 * self.nat0,
                                              int(maxit), float(tol), C.byref(it), C.byref(te), C.byref(nc))
        if check:
            self._check(f"solve_device_{name}", rc)
        return it.value, te.value, nc.value, rc

    def apply_batch(self, cwhat, nrhs, x_ptr, y_ptr, adia_ptr, aoff_ptr, op_per_rhs=True, stream=None):
        cw = ord(cwhat) if isinstance(cwhat, str) else int(cwhat)
        return self._check("apply_batch", self.lib.ddscat_hip_apply_batch(
            cw, int(nrhs), C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(adia_ptr), C.c_void_p(aoff_ptr),
            1 if op_per_rhs else 0, C.c_void_p(stream) if stream else None))

    def evalq_batch(self, nrhs, e_ptr, p_ptr, adia_ptr, aoff_ptr, ak, e02=1.0, op_per_rhs=True, check=True):
        """Cabs, Cext, Cpha per RHS ([nrhs][3] float64) from device vectors."""
        ak = np.ascontiguousarray(np.broadcast_to(np.asarray(ak, np.float32).reshape(-1, 3), (nrhs, 3)))
        out = np.zeros((nrhs, 3), np.float64)
        rc = self.lib.ddscat_hip_evalq_batch(int(nrhs), C.c_void_p(e_ptr), C.c_void_p(p_ptr), C.c_void_p(adia_ptr),
                                             C.c_void_p(aoff_ptr), 1 if op_per_rhs else 0, _p(ak), float(e02),
                                             _p(out))
        if check:
            self._check("evalq_batch", rc)
        return out

    def cabs_adjoint(self, name, nrhs, p_ptr, adia_ptr, aoff_ptr, ak, g_adia_ptr, g_aoff_ptr=None, e02=1.0,
                     op_per_rhs=True, maxit=300, tol=1e-5, check=True):
        """Writes dCabs/d(adia) (and d(aoff)) as complex64 device arrays; returns (itno, tole, rc)."""
        ak = np.ascontiguousarray(np.broadcast_to(np.asarray(ak, np.float32).reshape(-1, 3), (nrhs, 3)))
        it = np.zeros(nrhs, np.int32)
        te = np.zeros(nrhs, np.float32)
        rc = self.lib.ddscat_hip_cabs_adjoint(SOLVER_IDS[name], int(nrhs), C.c_void_p(p_ptr), C.c_void_p(adia_ptr),
                                              C.c_void_p(aoff_ptr), 1 if op_per_rhs else 0, _p(ak), float(e02),
                                              int(maxit), float(tol), C.c_void_p(g_adia_ptr),
                                              C.c_void_p(g_aoff_ptr) if g_aoff_ptr else None, _p(it), _p(te))
        if check:
            self._check("cabs_adjoint", rc)
        return it, te, rc

    def setup_rhs(self, ak, e0, x0, eps_ptr, b_ptr, adia_ptr, aoff_ptr, dx=(1.0, 1.0, 1.0), eps_per_comp=False):
        """Device incident fields and LDR operators for RHS r = (ak[r], e0[r]) into
        compact [nrhs][3][NAT0] device arrays (eps: complex64 device array)."""
        ak = np.ascontiguousarray(np.asarray(ak, np.float32).reshape(-1, 3))
        nrhs = ak.shape[0]
        e0 = np.ascontiguousarray(np.asarray(e0, np.complex64).reshape(nrhs, 3))
        x0 = np.ascontiguousarray(np.asarray(x0, np.float32).reshape(3))
        dx = np.ascontiguousarray(np.asarray(dx, np.float32).reshape(3))
        return self._check("setup_rhs", self.lib.ddscat_hip_setup_rhs(
            nrhs, _p(ak), _p(e0), _p(x0), _p(dx), C.c_void_p(eps_ptr), 1 if eps_per_comp else 0,
            C.c_void_p(b_ptr), C.c_void_p(adia_ptr), C.c_void_p(aoff_ptr)))

    def solve_batch(self, name, nrhs, b_ptr, x_ptr, adia_ptr, aoff_ptr, op_per_rhs=True, maxit=300, tol=1e-5,
                    graph=True, check=True):
        """Batched solve on compact device vectors [nrhs][3][NAT0] (x in/out).
        Returns (itno[nrhs], tole[nrhs], ncompte[nrhs], rc)."""
        it = np.zeros(nrhs, np.int32)
        nc = np.zeros(nrhs, np.int32)
        te = np.zeros(nrhs, np.float32)
        rc = self.lib.ddscat_hip_solve_batch(SOLVER_IDS[name], int(nrhs), C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                             C.c_void_p(adia_ptr), C.c_void_p(aoff_ptr), 1 if op_per_rhs else 0,
                                             int(maxit), float(tol), _p(it), _p(te), _p(nc), 1 if graph else 0)
        if check:
            self._check(f"solve_batch_{name}", rc)
        return it, te, nc, rc

    # --------------------------------------------------------------- evalq
    def evalq(self, ak, e02=1.0, use_resident=True, e_c=None, p_c=None, adia_c=None, aoff_c=None, check=True):
        out = (C.c_double * 3)()
        ptrs = [None] * 4
        keep = []
        if not use_resident:
            n = 3 * self.nat0
            arrs = [_c64(a, n) for a in (e_c, p_c, adia_c, aoff_c)]
            keep = arrs
            ptrs = [_p(a) for a in arrs]
        rc = self.lib.ddscat_cuda_evalq_f32(1 if use_resident else 0, *ptrs, float(ak[0]), float(ak[1]),
                                            float(ak[2]), float(e02), C.byref(out, 0), C.byref(out, 8),
                                            C.byref(out, 16))
        del keep
        if check:
            self._check("evalq", rc)
        return out[0], out[1], out[2], rc

    # ---------------------------------------------------------------- scat
    def scat(self, ak, dx, x0, e01, etasca, e02=1.0, aks=None, em1=None, em2=None, jpbc=0, check=True,
             noalloc=False):
        """ddscat_cuda_scat_f32, or its allocation-free variant (noalloc=True)."""
        ak = np.ascontiguousarray(ak, dtype=np.float32)
        dx = np.ascontiguousarray(dx, dtype=np.float32)
        x0 = np.ascontiguousarray(x0, dtype=np.float32)
        e01 = _c64(e01, 3)
        ndir = 0 if aks is None else len(aks)
        if ndir:
            aks = np.ascontiguousarray(aks, dtype=np.float32).reshape(ndir, 3)
            em1 = np.ascontiguousarray(em1, dtype=np.float32).reshape(ndir, 3)
            em2 = np.ascontiguousarray(em2, dtype=np.float32).reshape(ndir, 3)
            f1 = np.zeros(ndir, np.complex64)
            f2 = np.zeros(ndir, np.complex64)
        else:
            aks = em1 = em2 = f1 = f2 = np.zeros(3, np.float32)
        cb, cs, cg2 = C.c_double(), C.c_double(), C.c_double()
        cg = (C.c_double * 3)()
        navg = C.c_int(0)
        fn = self.lib.ddscat_hip_scat_f32_noalloc if noalloc else self.lib.ddscat_cuda_scat_f32
        rc = fn(jpbc, ndir, _p(ak), _p(aks), _p(dx), _p(em1), _p(em2), float(e02), float(etasca), _p(e01), _p(x0),
                C.byref(cb), C.byref(cs), cg, C.byref(cg2), C.byref(navg), _p(f1), _p(f2))
        if check:
            self._check("scat", rc)
        res = dict(cbksca=cb.value, csca=cs.value, cscag=np.array(cg[:]), cscag2=cg2.value, navg=navg.value, rc=rc)
        if ndir:
            res["f1"], res["f2"] = f1, f2
        return res

    def torque(self, ak, e01, etasca, dx=(1.0, 1.0, 1.0), x0=(0.0, 0.0, 0.0), e02=None, check=True):
        """Torque cross sections (ext, rad, total [3] each, navg) of the resident
        solution; e02 defaults to |e01|^2."""
        ak = np.ascontiguousarray(ak, dtype=np.float32)
        dx = np.ascontiguousarray(dx, dtype=np.float32)
        x0 = np.ascontiguousarray(x0, dtype=np.float32)
        e01 = _c64(e01, 3)
        e02 = float(np.sum(np.abs(e01.astype(np.complex128)) ** 2)) if e02 is None else float(e02)
        ce = np.zeros(3)
        cr = np.zeros(3)
        navg = C.c_int(0)
        rc = self.lib.ddscat_hip_torque(_p(ak), _p(e01), e02, float(etasca), _p(dx), _p(x0), _p(ce), _p(cr),
                                        C.byref(navg))
        if check:
            self._check("torque", rc)
        return dict(ext=ce, rad=cr, total=ce + cr, navg=navg.value, rc=rc)

    # ----------------------------------------------------------- extensions
    def set_backend(self, backend):
        return self._check("set_backend", self.lib.ddscat_hip_set_backend(int(backend)))

    def set_precision(self, prec):
        return self._check("set_precision", self.lib.ddscat_hip_set_precision(int(prec)))

    def set_fft_kernels(self, generic):
        """0: register-resident FFTs where available (default); 1: LDS-staged everywhere."""
        return self._check("set_fft_kernels", self.lib.ddscat_hip_set_fft_kernels(int(generic)))

    def nearfield(self, p_ptr, out_ptr, einc_ptr=None):
        """E_inc + E_sca (or E_sca) on every lattice site, full [3][NAT] device array."""
        return self._check("nearfield", self.lib.ddscat_hip_nearfield(
            C.c_void_p(p_ptr), C.c_void_p(einc_ptr) if einc_ptr else None, C.c_void_p(out_ptr)))

    def scat_matrix(self, p_ptr, npair, ak, e0, aks, em1, em2, x0, dx=(1.0, 1.0, 1.0), e02=1.0,
                    mueller=True):
        """Amplitudes f11 f21 f12 f22 [npair, ndir, 4] and Mueller matrices
        [npair, ndir, 4, 4] of incident-polarization pairs whose compact
        polarizations sit at the device pointer p_ptr ([2*npair][3][NAT0])."""
        ak = np.ascontiguousarray(np.asarray(ak, np.float32).reshape(npair, 3))
        e0 = np.ascontiguousarray(np.asarray(e0, np.complex64).reshape(npair, 2, 3))
        aks = np.ascontiguousarray(np.asarray(aks, np.float32).reshape(npair, -1, 3))
        ndir = aks.shape[1]
        em1 = np.ascontiguousarray(np.asarray(em1, np.float32).reshape(npair, ndir, 3))
        em2 = np.ascontiguousarray(np.asarray(em2, np.float32).reshape(npair, ndir, 3))
        dxa = np.asarray(dx, np.float32)
        x0a = np.asarray(x0, np.float32)
        f = np.empty((npair, ndir, 4), np.complex64)
        M = np.empty((npair, ndir, 4, 4), np.float64) if mueller else None
        self._check("scat_matrix", self.lib.ddscat_hip_scat_matrix(
            int(npair), C.c_void_p(p_ptr), int(ndir), _p(ak), _p(e0), _p(aks), _p(em1), _p(em2), _p(dxa), _p(x0a),
            float(e02), _p(f), _p(M) if mueller else None))
        return f, M

    def set_pass_kernel(self, pass_, kind, check=True):
        rc = self.lib.ddscat_hip_set_pass_kernel(int(pass_), int(kind))
        if check:
            self._check("set_pass_kernel", rc)
        return rc

    def force_restart(self):
        """The next reliable-residual check of a single-RHS solve restarts (test hook)."""
        return self._check("force_restart", self.lib.ddscat_hip_force_restart())

    def poison_workspace(self):
        """Fill the solver workspaces (single-RHS, and batched when allocated)
        with NaN bit patterns (test hook)."""
        return self._check("poison_workspace", self.lib.ddscat_hip_poison_workspace())

    def solve_log(self, rhs=0) -> dict:
        """Convergence record of RHS `rhs` of the last solve: restarts, the
        recursive residual after each iteration, and the reliable checks as
        rows (iteration, true residual, gap, action: 0 none, 1 restart, 2
        residual replaced)."""
        nrst = C.c_int(0)
        nck = C.c_int(0)
        n = self.lib.ddscat_hip_solve_log(int(rhs), C.byref(nrst), None, 0, None, 0, C.byref(nck))
        if n < 0:
            raise IndexError(f"no RHS {rhs} in the last solve")
        resid = np.zeros(n, np.float64)
        checks = np.zeros((nck.value, 4), np.float64)
        self.lib.ddscat_hip_solve_log(int(rhs), C.byref(nrst), _p(resid), n, _p(checks), nck.value, C.byref(nck))
        return dict(restarts=nrst.value, resid=resid, checks=checks)

    def set_polarizability(self, prescription):
        """LDR prescription of setup_rhs: "LATTDR" (default) or "GKDLDR"."""
        code = {"LATTDR": 0, "GKDLDR": 1}.get(prescription, -1)
        return self._check("set_polarizability", self.lib.ddscat_hip_set_polarizability(code))

    def set_reliable_mode(self, mode):
        """QMRCCG's response to a failed reliable check: "replace" (default) or "restart" (the reference's)."""
        code = {"replace": 0, "restart": 1}.get(mode, -1)
        return self._check("set_reliable_mode", self.lib.ddscat_hip_set_reliable_mode(code))

    def select_target(self, slot):
        """Make target slot `slot` (0 = the reference ABI's) the one every call acts on."""
        return self._check("select_target", self.lib.ddscat_hip_select_target(int(slot)))

    def solve_targets(self, name, slots, nrhs, b_ptrs, x_ptrs, adia_ptrs, aoff_ptrs, op_per_rhs=True, maxit=300,
                      tol=1e-5, graph=True, check=True):
        """Batched solves of several prepared targets at once (one stream and
        host thread each).  Returns (itno, tole, ncompte) per target and rc."""
        n = len(slots)
        sl = np.ascontiguousarray(slots, np.int32)
        nr = np.ascontiguousarray(nrhs, np.int32)
        arr = lambda ptrs: (C.c_void_p * n)(*[C.c_void_p(int(v)) for v in ptrs])
        tot = int(nr.sum())
        it = np.zeros(tot, np.int32)
        nc = np.zeros(tot, np.int32)
        te = np.zeros(tot, np.float32)
        rc = self.lib.ddscat_hip_solve_targets(SOLVER_IDS[name], n, _p(sl), _p(nr), arr(b_ptrs), arr(x_ptrs),
                                               arr(adia_ptrs), arr(aoff_ptrs), 1 if op_per_rhs else 0, int(maxit),
                                               float(tol), _p(it), _p(te), _p(nc), 1 if graph else 0)
        if check:
            self._check(f"solve_targets_{name}", rc)
        cut = np.cumsum(np.concatenate([[0], nr]))
        per = [(it[cut[t]:cut[t + 1]], te[cut[t]:cut[t + 1]], nc[cut[t]:cut[t + 1]]) for t in range(n)]
        return per, rc

    def set_autotune(self, on):
        return self._check("set_autotune", self.lib.ddscat_hip_set_autotune(1 if on else 0))

    def apply_device(self, cwhat, x_ptr, y_ptr, adia_ptr, aoff_ptr, layout=0, stream=None):
        cw = ord(cwhat) if isinstance(cwhat, str) else int(cwhat)
        return self._check("apply_device", self.lib.ddscat_hip_apply_device(
            cw, C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(adia_ptr), C.c_void_p(aoff_ptr), int(layout),
            C.c_void_p(stream) if stream else None))

    def info(self) -> dict:
        buf = (C.c_longlong * len(INFO_FIELDS))()
        self.lib.ddscat_hip_info(buf, len(INFO_FIELDS))
        return dict(zip(INFO_FIELDS, buf[:]))

    def get_green(self):
        n = self.info()["green_count"]
        g = np.empty(n, dtype=np.complex64)
        self._check("get_green", self.lib.ddscat_hip_get_green(_p(g)))
        return g

    def time_matvec(self, cwhat, x_ptr, y_ptr, adia_ptr, aoff_ptr, layout=1, reps=10):
        """Mean per-pass and per-MATVEC milliseconds (HIP events on the library stream)."""
        out = (C.c_double * 6)()
        cw = ord(cwhat) if isinstance(cwhat, str) else int(cwhat)
        self._check("time_matvec", self.lib.ddscat_hip_time_matvec(
            cw, C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(adia_ptr), C.c_void_p(aoff_ptr), int(layout),
            int(reps), out))
        return list(out)

    def stream(self):
        return self.lib.ddscat_hip_stream()

    def sync(self):
        return self._check("sync", self.lib.ddscat_hip_sync())


def exported_symbols(path) -> set[str]:
    """Names the shared object exports (dlsym probe; no GPU call)."""
    lib = C.CDLL(str(path), mode=C.RTLD_LOCAL)
    found = set()
    for name in REFERENCE_SYMBOLS + EXTENSION_SYMBOLS:
        try:
            getattr(lib, name)
            found.add(name)
        except AttributeError:
            pass
    return found
