"""ctypes binding of libsparcgpu.so, mirroring the reference entry points.

`Context` owns one grid on one GPU.  Its methods carry the reference's names
and argument meaning (column-major blocks, Nd x ncol, FP64 host arrays):

    Lap_vec_mult                  src/lapVecRoutines.c:37
    Lap_plus_diag_vec_mult        src/lapVecRoutines.c:306
    Vnl_vec_mult                  src/nlocVecRoutines.c:798
    Hamiltonian_vectors_mult      src/hamiltonianVecRoutines.c:45
    ChebyshevFiltering            src/eigenSolver.c:722
    DP_Project_Hamiltonian        src/eigenSolver.c:939
    DP_Solve_Generalized_EigenProblem   src/eigenSolver.c:1262
    DP_Subspace_Rotation          src/eigenSolver.c:1386
    AAR                           src/linearSolver.c:37
    Lanczos                       src/eigenSolver.c:1920

Errors raise SparcGPUError with the library's message (the reference exits).
The library is loaded from sparc_amd/lib; there is no CPU fallback: if it is
missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os
from pathlib import Path

import numpy as np

from .workloads import Grid, NonlocalSet

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libsparcgpu.so"

F64 = 0
F32 = 1
OK = 0
NOT_CONVERGED = 1


class SparcGPUError(RuntimeError):
    pass


_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_vp = C.c_void_p

class SCFParams(C.Structure):
    _fields_ = [("nstates", C.c_int), ("cheb_degree", C.c_int), ("rho_trigger", C.c_int),
                ("max_scf", C.c_int), ("min_scf", C.c_int), ("tol_scf", C.c_double),
                ("tol_poisson", C.c_double), ("maxit_poisson", C.c_int),
                ("tol_lanczos", C.c_double), ("beta", C.c_double), ("smear_type", C.c_int),
                ("nelec", C.c_double), ("occfac", C.c_double), ("mix_history", C.c_int),
                ("pulay_freq", C.c_int), ("mix_beta", C.c_double), ("mix_omega", C.c_double),
                ("kerker_ktf", C.c_double), ("kerker_thresh", C.c_double),
                ("tol_precond", C.c_double), ("Esc", C.c_double), ("xc", C.c_int),
                ("dtype", C.c_int),
                ("psi_seed", C.c_ulonglong), ("lanczos_seed", C.c_ulonglong),
                ("poisson_precond", C.c_void_p), ("poisson_precond_user", C.c_void_p),
                ("kerker_precond", C.c_void_p), ("kerker_precond_user", C.c_void_p),
                ("psi_out", C.c_void_p), ("vxc_out", C.c_void_p)]


class SCFResult(C.Structure):
    _fields_ = [("scf_iters", C.c_int), ("converged", C.c_int), ("Etot", C.c_double),
                ("Eband", C.c_double), ("Exc", C.c_double), ("E1", C.c_double),
                ("E2", C.c_double), ("E3", C.c_double), ("Entropy", C.c_double),
                ("Efermi", C.c_double), ("scf_error", C.c_double), ("mean_iter_ms", C.c_double),
                ("total_ms", C.c_double), ("poisson_iters", C.c_long),
                ("kerker_iters", C.c_long), ("poisson_fallbacks", C.c_int),
                ("kerker_fallbacks", C.c_int)]


class PspRadial(C.Structure):
    _fields_ = [("size", C.c_int), ("r", _dp), ("rvloc", _dp), ("rvloc_d", _dp), ("rhoc", _dp),
                ("rhoc_d", _dp), ("zion", C.c_double), ("vloc0", C.c_double)]


RES_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp)
PRECOND_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, _dp, _dp)

_SIGS = {
    "sgpu_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                              C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]),
    "sgpu_destroy": (C.c_int, [_vp]),
    "sgpu_last_error": (C.c_char_p, []),
    "sgpu_get_stencil_coeffs": (C.c_int, [_vp, _dp, _dp, _dp]),
    "sgpu_stencil_tile": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _ip]),
    "sgpu_stencil_kernel": (C.c_int, [_vp, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "sgpu_sync": (C.c_int, [_vp]),
    "sgpu_set_veff": (C.c_int, [_vp, _dp]),
    "sgpu_set_nonlocal": (C.c_int, [_vp, C.c_int, _ip, _ip, _lp, _ip, _lp, _dp, _ip, C.c_int, _dp,
                                    C.c_double]),
    "sgpu_block_create": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "sgpu_block_destroy": (C.c_int, [_vp]),
    "sgpu_block_ncol": (C.c_int, [_vp]),
    "sgpu_block_dtype": (C.c_int, [_vp]),
    "sgpu_block_ld": (C.c_long, [_vp]),
    "sgpu_block_upload": (C.c_int, [_vp, _dp, C.c_long, C.c_int]),
    "sgpu_block_download": (C.c_int, [_vp, _dp, C.c_long, C.c_int]),
    "sgpu_block_fill_uniform": (C.c_int, [_vp, C.c_ulonglong, C.c_double, C.c_double]),
    "sgpu_block_copy": (C.c_int, [_vp, _vp]),
    "sgpu_Lap_vec_mult": (C.c_int, [_vp, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int]),
    "sgpu_Lap_plus_diag_vec_mult": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                              _dp, C.c_int, _dp, C.c_int]),
    "sgpu_Vnl_vec_mult": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _dp, C.c_int]),
    "sgpu_Hamiltonian_vectors_mult": (C.c_int, [_vp, C.c_int, C.c_double, _dp, C.c_int, _dp,
                                                C.c_int]),
    "sgpu_ChebyshevFiltering": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int,
                                          C.c_double, C.c_double, C.c_double, C.c_int]),
    "sgpu_DP_Project_Hamiltonian": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp,
                                              C.c_int]),
    "sgpu_DP_Solve_Generalized_EigenProblem": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "sgpu_DP_Subspace_Rotation": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int]),
    "sgpu_AAR": (C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, _dp, _dp,
                           C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp]),
    "sgpu_lap_plus_diag_dev": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, _vp, _vp]),
    "sgpu_hamiltonian_dev": (C.c_int, [_vp, C.c_double, _vp, _vp]),
    "sgpu_hamiltonian_dev_noalloc": (C.c_int, [_vp, C.c_double, _vp, _vp]),
    "sgpu_chebyshev_filter_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_double,
                                            C.c_double]),
    "sgpu_project_dev": (C.c_int, [_vp, _vp, _vp, _dp, _dp]),
    "sgpu_rotate_dev": (C.c_int, [_vp, _vp, _dp, _vp]),
    "sgpu_lanczos": (C.c_int, [_vp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp, _ip]),
    "sgpu_chefsi_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp]),
    "sgpu_scf_run": (C.c_int, [_vp, C.POINTER(SCFParams), _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                               _dp, C.POINTER(SCFResult)]),
    "sgpu_unet_create": (C.c_int, [_vp, C.c_int, C.c_ulonglong, C.POINTER(_vp)]),
    "sgpu_unet_destroy": (None, [_vp]),
    "sgpu_unet_nparams": (C.c_int, [_vp]),
    "sgpu_unet_get_weights": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "sgpu_unet_set_weights": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "sgpu_unet_precond": (None, [_vp, C.c_int, C.c_double, _dp, _dp]),
    "sgpu_unet_apply": (C.c_int, [_vp, _dp, _dp]),
    "sgpu_unet_loss_grad": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp, C.POINTER(C.c_float)]),
    "sgpu_unet_harvest": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_int, _ip]),
    "sgpu_unet_set_shift": (C.c_int, [_vp, C.c_double]),
    "sgpu_nonlocal_integrals": (C.c_int, [_vp, _vp, _dp]),
    "sgpu_nonlocal_layout": (C.c_int, [_vp, _ip, C.POINTER(C.c_long)]),
    "sgpu_local_forces": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _ip, _ip, _ip, _dp, C.c_double, _dp,
                                    _dp, _dp, _dp, C.c_int, _dp, _dp]),
    "sgpu_unet_npairs": (C.c_int, [_vp]),
    "sgpu_unet_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_ulonglong, _dp]),
    "sgpu_timer_get": (C.c_int, [_vp, C.c_char_p, _dp, _ip]),
    "sgpu_timer_reset": (C.c_int, [_vp]),
    "sgpu_timer_enable": (C.c_int, [_vp, C.c_int]),
}


def exported_symbols() -> list[str]:
    return list(_SIGS)


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libsparcgpu.so (no GPU call is made by loading)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise SparcGPUError(f"{p} not built; run `python -m sparc_amd.build` (no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray(a) if a.ndim == 2 else np.ascontiguousarray(a)


def _col_block(x: np.ndarray, nd: int):
    """(array, ld, ncol) of a column-major Nd x ncol (or length-Nd) block."""
    x = np.asarray(x)
    if x.ndim == 1:
        return x, max(nd, x.shape[0]), 1
    return x, x.shape[0], x.shape[1]


class Block:
    """Device-resident Nd x ncol block (FP64 or FP32) owned by a Context."""

    def __init__(self, ctx: "Context", ncol: int, dtype: int = F64):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(ctx.lib.sgpu_block_create(ctx.h, ncol, dtype, C.byref(h)))
        self.h = h
        self.ncol = ncol
        self.dtype = dtype

    def upload(self, host: np.ndarray):
        a = np.asfortranarray(np.asarray(host, dtype=np.float64).reshape(self.ctx.nd, -1))
        self.ctx._check(self.ctx.lib.sgpu_block_upload(self.h, _ptr(a), a.shape[0], a.shape[1]))
        return self

    def download(self, ncol: int | None = None) -> np.ndarray:
        n = self.ncol if ncol is None else ncol
        out = np.zeros((self.ctx.nd, n), order="F")
        self.ctx._check(self.ctx.lib.sgpu_block_download(self.h, _ptr(out), self.ctx.nd, n))
        return out

    def fill_uniform(self, seed: int, lo: float, hi: float):
        self.ctx._check(self.ctx.lib.sgpu_block_fill_uniform(self.h, seed, lo, hi))
        return self

    def copy_from(self, other: "Block"):
        self.ctx._check(self.ctx.lib.sgpu_block_copy(other.h, self.h))
        return self

    def free(self):
        if self.h:
            self.ctx.lib.sgpu_block_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    def __init__(self, grid: Grid, device: int = 0, lib: C.CDLL | None = None):
        self.lib = lib or load()
        self.grid = grid
        self.nd = grid.nd
        h = C.c_void_p()
        self._check(self.lib.sgpu_create(C.byref(h), device, grid.nx, grid.ny, grid.nz, grid.hx,
                                         grid.hy, grid.hz, grid.bcx, grid.bcy, grid.bcz, grid.order))
        self.h = h
        self._keep = []
        self._nets = weakref.WeakSet()  # closed with the context (C: orphaned if left open)
        self._n_alpha = 0

    # -- plumbing
    def _check(self, st: int) -> int:
        if st < 0:
            raise SparcGPUError(self.lib.sgpu_last_error().decode(errors="replace"))
        return st

    def close(self):
        if getattr(self, "h", None):
            for net in list(getattr(self, "_nets", ())):
                net.close()
            self.lib.sgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self._check(self.lib.sgpu_sync(self.h))

    def stencil_coeffs(self):
        n = self.grid.order // 2 + 1
        a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(self.lib.sgpu_get_stencil_coeffs(self.h, _ptr(a), _ptr(b), _ptr(c)))
        return a, b, c

    def stencil_tile(self, dtype: int, ncol: int):
        """(tx, ty, ry) of the stencil tile used for ncol columns of dtype, or
        None before a launch of that shape has chosen one."""
        t = [C.c_int() for _ in range(3)]
        self._check(self.lib.sgpu_stencil_tile(self.h, dtype, ncol, *[C.byref(v) for v in t]))
        return None if t[0].value == 0 else tuple(v.value for v in t)

    def stencil_kernel(self, dtype: int, ncol: int):
        """Name of the stencil kernel launched for ncol columns of dtype (as
        the profiler lists it, without spaces), or None before a launch of
        that shape has chosen one."""
        buf = C.create_string_buffer(96)
        self._check(self.lib.sgpu_stencil_kernel(self.h, dtype, ncol, buf, 96))
        return buf.value.decode() or None

    def set_veff(self, veff: np.ndarray):
        v = _f64(veff).ravel()
        assert v.size == self.nd
        self._check(self.lib.sgpu_set_veff(self.h, _ptr(v)))

    def set_nonlocal(self, nl: NonlocalSet | None):
        self._n_alpha = 0 if nl is None or nl.n_influence == 0 else int(nl.n_alpha)
        if nl is None or nl.n_influence == 0:
            z = np.zeros(1, np.int32)
            self._check(self.lib.sgpu_set_nonlocal(self.h, 0, None, None, None, None, None, None,
                                                   None, 0, None, 0.0))
            return
        arrs = [np.ascontiguousarray(nl.ndc, np.int32), np.ascontiguousarray(nl.nproj, np.int32),
                np.ascontiguousarray(nl.pos_off, np.int64), np.ascontiguousarray(nl.grid_pos, np.int32),
                np.ascontiguousarray(nl.chi_off, np.int64), np.ascontiguousarray(nl.chi, np.float64),
                np.ascontiguousarray(nl.alpha_off, np.int32), np.ascontiguousarray(nl.gamma, np.float64)]
        ndc, nproj, pos_off, grid_pos, chi_off, chi, alpha_off, gamma = arrs
        self._check(self.lib.sgpu_set_nonlocal(
            self.h, nl.n_influence, ndc.ctypes.data_as(_ip), nproj.ctypes.data_as(_ip),
            pos_off.ctypes.data_as(_lp), grid_pos.ctypes.data_as(_ip), chi_off.ctypes.data_as(_lp),
            _ptr(chi), alpha_off.ctypes.data_as(_ip), int(nl.n_alpha), _ptr(gamma), float(nl.dV)))

    def block(self, ncol: int, dtype: int = F64) -> Block:
        return Block(self, ncol, dtype)

    # -- reference entry points (host arrays)
    def Lap_vec_mult(self, x, c: float = 0.0):
        x, ld, ncol = _col_block(_f64(x), self.nd)
        y = np.zeros_like(x)
        self._check(self.lib.sgpu_Lap_vec_mult(self.h, ncol, c, _ptr(x), ld, _ptr(y), ld))
        return y

    def Lap_plus_diag_vec_mult(self, x, a: float, b: float, c: float, v=None):
        x, ld, ncol = _col_block(_f64(x), self.nd)
        vv = None if v is None else _f64(v).ravel()
        y = np.zeros_like(x)
        self._check(self.lib.sgpu_Lap_plus_diag_vec_mult(self.h, ncol, a, b, c, _ptr(vv), _ptr(x), ld,
                                                         _ptr(y), ld))
        return y

    def Vnl_vec_mult(self, x, Hx=None):
        x, ld, ncol = _col_block(_f64(x), self.nd)
        out = np.zeros_like(x) if Hx is None else _f64(Hx).copy(order="K")
        self._check(self.lib.sgpu_Vnl_vec_mult(self.h, ncol, _ptr(x), ld, _ptr(out), ld))
        return out

    def Hamiltonian_vectors_mult(self, x, c: float = 0.0):
        x, ld, ncol = _col_block(_f64(x), self.nd)
        y = np.zeros_like(x)
        self._check(self.lib.sgpu_Hamiltonian_vectors_mult(self.h, ncol, c, _ptr(x), ld, _ptr(y), ld))
        return y

    def ChebyshevFiltering(self, X, m: int, a: float, b: float, a0: float, dtype: int = F64):
        """Returns (Y, X_clobbered) like the reference (Y filtered, X = Y_{m-1})."""
        X = np.asfortranarray(np.array(X, dtype=np.float64, copy=True).reshape(self.nd, -1))
        Y = np.zeros_like(X)
        ld, ncol = X.shape
        self._check(self.lib.sgpu_ChebyshevFiltering(self.h, _ptr(X), ld, _ptr(Y), ld, ncol, m, a, b,
                                                     a0, dtype))
        return Y, X

    def DP_Project_Hamiltonian(self, Y, dtype: int = F64):
        Y = np.asfortranarray(np.asarray(Y, np.float64).reshape(self.nd, -1))
        ld, ncol = Y.shape
        HY = np.zeros_like(Y)
        Hp = np.zeros((ncol, ncol), order="F")
        Mp = np.zeros((ncol, ncol), order="F")
        self._check(self.lib.sgpu_DP_Project_Hamiltonian(self.h, _ptr(Y), ld, _ptr(HY), ld, ncol,
                                                         _ptr(Hp), _ptr(Mp), dtype))
        return Hp, Mp, HY

    def DP_Solve_Generalized_EigenProblem(self, Hp, Mp):
        H = np.asfortranarray(np.array(Hp, np.float64, copy=True))
        M = np.asfortranarray(np.array(Mp, np.float64, copy=True))
        n = H.shape[0]
        lam = np.zeros(n)
        self._check(self.lib.sgpu_DP_Solve_Generalized_EigenProblem(self.h, n, _ptr(H), _ptr(M),
                                                                    _ptr(lam)))
        return lam, H

    def DP_Subspace_Rotation(self, Y, Q, dtype: int = F64):
        Y = np.asfortranarray(np.asarray(Y, np.float64).reshape(self.nd, -1))
        Q = np.asfortranarray(np.asarray(Q, np.float64))
        ld, ns = Y.shape
        out = np.zeros_like(Y)
        self._check(self.lib.sgpu_DP_Subspace_Rotation(self.h, _ptr(Y), ld, _ptr(Q), ns, _ptr(out), ld,
                                                       dtype))
        return out

    def AAR(self, x0, b, c: float = 0.0, omega: float = 0.6, beta: float = 0.6, m: int = 7,
            p: int = 6, tol: float = 1e-8, max_iter: int = 1000, res=None, precond=None, user=None):
        """Returns (x, iterations, relres, converged)."""
        x = np.array(x0, np.float64, copy=True).ravel()
        bb = _f64(b).ravel()
        it = C.c_int(0)
        rr = C.c_double(0.0)
        res_c = RES_FN(res) if res is not None else None
        pre_c = PRECOND_FN(precond) if precond is not None else None
        st = self._check(self.lib.sgpu_AAR(
            self.h, C.cast(res_c, C.c_void_p) if res_c else None,
            C.cast(pre_c, C.c_void_p) if pre_c else None, user, c, self.nd, _ptr(x), _ptr(bb), omega,
            beta, m, p, tol, max_iter, C.byref(it), C.byref(rr)))
        return x, it.value, rr.value, st == OK

    def Lanczos(self, x0, tol_min: float, tol_max: float, maxit: int):
        x = _f64(x0).ravel()
        emin, emax, it = C.c_double(), C.c_double(), C.c_int()
        self._check(self.lib.sgpu_lanczos(self.h, _ptr(x), tol_min, tol_max, maxit, C.byref(emin),
                                          C.byref(emax), C.byref(it)))
        return emin.value, emax.value, it.value

    # -- device-resident operations
    def lap_plus_diag_dev(self, a: float, b: float, c: float, x: Block, y: Block):
        self._check(self.lib.sgpu_lap_plus_diag_dev(self.h, a, b, c, x.h, y.h))

    def hamiltonian_dev(self, c: float, x: Block, Hx: Block):
        self._check(self.lib.sgpu_hamiltonian_dev(self.h, c, x.h, Hx.h))

    def hamiltonian_dev_noalloc(self, c: float, x: Block, Hx: Block):
        """sgpu_hamiltonian_dev without any device allocation (the library's
        fixed per-GPU scratch array holds the nonlocal partials)."""
        self._check(self.lib.sgpu_hamiltonian_dev_noalloc(self.h, c, x.h, Hx.h))

    def chebyshev_filter_dev(self, X: Block, Y: Block, m: int, a: float, b: float, a0: float):
        self._check(self.lib.sgpu_chebyshev_filter_dev(self.h, X.h, Y.h, m, a, b, a0))

    def project_dev(self, Y: Block, HY: Block):
        n = Y.ncol
        Hp = np.zeros((n, n), order="F")
        Mp = np.zeros((n, n), order="F")
        self._check(self.lib.sgpu_project_dev(self.h, Y.h, HY.h, _ptr(Hp), _ptr(Mp)))
        return Hp, Mp

    def rotate_dev(self, Y: Block, Q, out: Block):
        Q = np.asfortranarray(np.asarray(Q, np.float64))
        self._check(self.lib.sgpu_rotate_dev(self.h, Y.h, _ptr(Q), out.h))

    def chefsi_dev(self, X: Block, Y: Block, m: int, a: float, b: float, a0: float):
        lam = np.zeros(X.ncol)
        self._check(self.lib.sgpu_chefsi_dev(self.h, X.h, Y.h, m, a, b, a0, _ptr(lam)))
        return lam

    # -- SCF
    def scf_run(self, b, rho0, params: dict, rho_core=None, poisson_net: "UNet | None" = None,
                kerker_net: "UNet | None" = None, keep_psi: bool = False):
        """Density-mixing SCF (sgpu_scf_run).  Returns a dict with the energies,
        eigenvalues, final density / potential and per-iteration history.
        poisson_net / kerker_net: a trained U-Net of this context as the
        preconditioner of the Poisson / Kerker AAR solves (None: Jacobi, as
        the reference); one network may serve both.  keep_psi: also return
        the final orbitals as a device Block ("psi") and Vxc ("vxc"), the
        inputs of the forces."""
        p = SCFParams(**params)
        psi = vxc = None
        if keep_psi:
            psi = self.block(p.nstates, p.dtype)
            vxc = np.zeros(self.nd)
            p.psi_out = psi.h.value
            p.vxc_out = vxc.ctypes.data
        pre = C.cast(self.lib.sgpu_unet_precond, C.c_void_p).value
        for net, fn, user in ((poisson_net, "poisson_precond", "poisson_precond_user"),
                              (kerker_net, "kerker_precond", "kerker_precond_user")):
            if net is None:
                continue
            if net.ctx is not self:
                raise ValueError("scf_run: the U-Net belongs to another context")
            setattr(p, fn, pre)
            setattr(p, user, net.h.value)
        bb = _f64(b).ravel()
        r0 = _f64(rho0).ravel()
        rc = None if rho_core is None else _f64(rho_core).ravel()
        ns, mx = p.nstates, p.max_scf
        lam = np.zeros(ns)
        rho = np.zeros(self.nd)
        phi = np.zeros(self.nd)
        ims, ierr, iet = np.zeros(mx), np.zeros(mx), np.zeros(mx)
        res = SCFResult()
        st = self._check(self.lib.sgpu_scf_run(self.h, C.byref(p), _ptr(bb), _ptr(r0), _ptr(rc),
                                               _ptr(lam),
                                               _ptr(rho), _ptr(phi), _ptr(ims), _ptr(ierr),
                                               _ptr(iet), C.byref(res)))
        k = res.scf_iters
        out = {f: getattr(res, f) for f, _ in SCFResult._fields_}
        out.update(lambda_=lam, rho=rho, phi=phi, iter_ms=ims[:k], iter_err=ierr[:k],
                   iter_etot=iet[:k], status=st)
        if keep_psi:
            out.update(psi=psi, vxc=vxc)
        return out

    def local_forces(self, setup, boxes, phi, vxc, ref_cutoff: float = 0.5):
        """(F^loc, F^xc) [n_atom, 3] (sgpu_local_forces); boxes = (type,
        atom, box, centre) arrays per image atom (sparc_amd.forces.image_boxes)."""
        types, atoms, box, centre = boxes
        keep = []
        tabs = (PspRadial * len(setup.types))()
        for i, t in enumerate(setup.types):
            psp = t.psp
            arrs = [np.ascontiguousarray(psp.r, np.float64), np.ascontiguousarray(psp.rVloc, np.float64),
                    np.ascontiguousarray(psp.spline("rVloc"), np.float64)]
            nlcc = bool(np.any(psp.rho_c_table))
            if nlcc:
                arrs += [np.ascontiguousarray(psp.rho_c_table, np.float64),
                         np.ascontiguousarray(psp.spline("rho_c_table"), np.float64)]
            keep += arrs
            tabs[i].size = int(psp.r.size)
            tabs[i].r, tabs[i].rvloc, tabs[i].rvloc_d = (_ptr(a) for a in arrs[:3])
            tabs[i].rhoc = _ptr(arrs[3]) if nlcc else None
            tabs[i].rhoc_d = _ptr(arrs[4]) if nlcc else None
            tabs[i].zion = float(psp.zion)
            tabs[i].vloc0 = float(psp.Vloc_0)
        types = np.ascontiguousarray(types, np.int32); atoms = np.ascontiguousarray(atoms, np.int32)
        box = np.ascontiguousarray(box, np.int32); centre = np.ascontiguousarray(centre, np.float64)
        bsum = _f64(np.asarray(setup.b) + np.asarray(setup.b_ref)).ravel()
        n_atom = int(setup.n_atom)
        floc, fxc = np.zeros((n_atom, 3)), np.zeros((n_atom, 3))
        self._check(self.lib.sgpu_local_forces(
            self.h, len(setup.types), C.cast(tabs, C.c_void_p), int(types.size),
            types.ctypes.data_as(_ip), atoms.ctypes.data_as(_ip), box.ctypes.data_as(_ip),
            _ptr(centre), float(ref_cutoff), _ptr(_f64(phi).ravel()), _ptr(_f64(setup.Vc).ravel()),
            _ptr(bsum), _ptr(_f64(vxc).ravel()), n_atom, _ptr(floc), _ptr(fxc)))
        return floc, fxc

    def nonlocal_layout(self):
        """(mode, stored): 0 bricks / 1 spheres, and the projector values stored
        per copy with zero padding (sgpu_nonlocal_layout)."""
        mode, stored = C.c_int(), C.c_long()
        self._check(self.lib.sgpu_nonlocal_layout(self.h, C.byref(mode), C.byref(stored)))
        return mode.value, stored.value

    def nonlocal_integrals(self, psi: "Block") -> np.ndarray:
        """[4, n_alpha, ncol]: dV <Chi, psi> and <Chi, d psi / dx_d>
        (sgpu_nonlocal_integrals)."""
        n_alpha = self._n_alpha
        out = np.zeros((4, n_alpha, psi.ncol))
        if n_alpha:
            self._check(self.lib.sgpu_nonlocal_integrals(self.h, psi.h, _ptr(out)))
        return out

    # -- learned preconditioner
    def unet(self, seed: int = 1) -> "UNet":
        return UNet(self, seed)

    def AAR_unet(self, net: "UNet", x0, b, tol: float = 1e-8, max_iter: int = 3000):
        """AAR with the U-Net as precond_fun (the reference's Poisson AAR
        parameters 0.6 / 0.6 / 7 / 6).  Returns (x, iterations, relres)."""
        x = np.array(x0, np.float64, copy=True).ravel()
        bb = _f64(b).ravel()
        it, rr = C.c_int(), C.c_double()
        pre = C.cast(self.lib.sgpu_unet_precond, C.c_void_p)
        st = self._check(self.lib.sgpu_AAR(self.h, None, pre, net.h, 0.0, self.nd, _ptr(x), _ptr(bb),
                                           0.6, 0.6, 7, 6, tol, max_iter, C.byref(it),
                                           C.byref(rr)))
        return x, it.value, rr.value, st

    # -- timers
    def timer_enable(self, on: bool = True):
        self._check(self.lib.sgpu_timer_enable(self.h, 1 if on else 0))

    def timer_reset(self):
        self._check(self.lib.sgpu_timer_reset(self.h))

    def timer(self, name: str):
        t, n = C.c_double(), C.c_int()
        self._check(self.lib.sgpu_timer_get(self.h, name.encode(), C.byref(t), C.byref(n)))
        return t.value, n.value


def train_multigrid(nets, steps: int, batch: int = 4, lr: float = 2e-3, seed: int = 7,
                    block: int = 50):
    """Adam over pairs harvested on several grids (one network per grid, each
    holding its own pairs): blocks of `block` steps on each network in turn,
    the weights carried from one network to the next (each keeps its own Adam
    moments).  Returns the loss of every step in order and the final weights,
    which every network then holds."""
    w = nets[0].get_weights()
    hist, done, k = [], 0, 0
    while done < steps:
        n = min(block, steps - done)
        net = nets[k % len(nets)]
        net.set_weights(w)
        hist.append(net.train(n, batch=batch, lr=lr, seed=seed + k))
        w = net.get_weights()
        done += n
        k += 1
    for net in nets:
        net.set_weights(w)
    return np.concatenate(hist), w


class UNet:
    """The learned AAR preconditioner (sgpu_unet_*) bound to one context."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        ctx._check(self.lib.sgpu_unet_create(ctx.h, 16, seed, C.byref(h)))
        self.h = h
        ctx._nets.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.sgpu_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nparams(self) -> int:
        return self.lib.sgpu_unet_nparams(self.h)

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.nparams, np.float32)
        self.ctx._check(self.lib.sgpu_unet_get_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_float))))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nparams
        self.ctx._check(self.lib.sgpu_unet_set_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_float))))

    def apply(self, r) -> np.ndarray:
        r = _f64(r).ravel()
        f = np.zeros_like(r)
        self.ctx._check(self.lib.sgpu_unet_apply(self.h, _ptr(r), _ptr(f)))
        return f

    def loss_grad(self, R, E):
        R = np.ascontiguousarray(R, np.float64)
        E = np.ascontiguousarray(E, np.float64)
        npair = R.shape[0]
        loss = C.c_double()
        g = np.zeros(self.nparams, np.float32)
        self.ctx._check(self.lib.sgpu_unet_loss_grad(self.h, _ptr(R), _ptr(E), npair, C.byref(loss),
                                                     g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def set_shift(self, c: float):
        """Operator -(Lap + c) of harvesting and training (0 Poisson, -kTF^2 Kerker)."""
        self.ctx._check(self.lib.sgpu_unet_set_shift(self.h, c))

    def harvest(self, b, x0, tol=1e-8, max_iter=3000, stride=2, max_pairs=64) -> int:
        added = C.c_int()
        self.ctx._check(self.lib.sgpu_unet_harvest(self.h, _ptr(_f64(b).ravel()),
                                                   _ptr(_f64(x0).ravel()), tol, max_iter, stride,
                                                   max_pairs, C.byref(added)))
        return added.value

    @property
    def npairs(self) -> int:
        return self.lib.sgpu_unet_npairs(self.h)

    def train(self, steps: int, batch: int = 4, lr: float = 1e-3, seed: int = 7) -> np.ndarray:
        hist = np.zeros(max(steps, 1))
        self.ctx._check(self.lib.sgpu_unet_train(self.h, steps, batch, lr, seed, _ptr(hist)))
        return hist[:steps]
