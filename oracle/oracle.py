"""Python access to the FP64 CPU oracle (liboracle.so) and, where built, to the
reference's own stencil (oracle/_ref/libsparcref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product
(sparc_amd/, libsparcgpu.so) never imports this module.

Besides the C restatement this module holds closed-form / library references
used to pin it (FFT solution of the periodic Poisson problem, the Fourier
symbol of the stencil, LAPACK dsygvd through scipy).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libsparcref.so"
MAX_FDN = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("bcx", C.c_int), ("bcy", C.c_int), ("bcz", C.c_int), ("fdn", C.c_int),
                ("d2x", C.c_double * (MAX_FDN + 1)), ("d2y", C.c_double * (MAX_FDN + 1)),
                ("d2z", C.c_double * (MAX_FDN + 1))]


class OrcNloc(C.Structure):
    _fields_ = [("n_influence", C.c_int), ("ndc", _ip), ("nproj", _ip), ("pos_off", _lp),
                ("grid_pos", _ip), ("chi_off", _lp), ("chi", _dp), ("alpha_off", _ip),
                ("n_alpha", C.c_int), ("gamma", _dp), ("dV", C.c_double)]


RES_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp)
PRE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, _dp, _dp)

_lib = None


def build():
    """Compile liboracle.so (and the reference stencil where /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if Path("/root/reference/src/lapVecRoutines.c").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_grid_init.argtypes = [C.POINTER(OrcGrid), C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_fd_d2_weights.argtypes = [C.c_int, _dp]
        L.orc_max_eig_mhalf_lap.argtypes = [C.POINTER(OrcGrid)]
        L.orc_max_eig_mhalf_lap.restype = C.c_double
        L.orc_mesh2chebdegree.argtypes = [C.c_double]
        L.orc_lap_plus_diag.argtypes = [C.POINTER(OrcGrid), C.c_int, C.c_double, C.c_double,
                                        C.c_double, _dp, _dp, C.c_int, _dp, C.c_int]
        L.orc_vnl_vec_mult.argtypes = [C.POINTER(OrcNloc), C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.orc_hamiltonian.argtypes = [C.POINTER(OrcGrid), _dp, C.POINTER(OrcNloc), C.c_int,
                                      C.c_double, _dp, C.c_int, _dp, C.c_int]
        L.orc_chebyshev_filter.argtypes = [C.POINTER(OrcGrid), _dp, C.POINTER(OrcNloc), _dp,
                                           C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double]
        L.orc_lanczos.argtypes = [C.POINTER(OrcGrid), _dp, C.POINTER(OrcNloc), _dp, C.c_double,
                                  C.c_double, C.c_int, _dp, _dp]
        L.orc_tridiag_minmax.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_project_hamiltonian.argtypes = [C.POINTER(OrcGrid), _dp, C.POINTER(OrcNloc), _dp,
                                              C.c_int, _dp, C.c_int, C.c_int, _dp, _dp]
        L.orc_subspace_rotation.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int]
        L.orc_sygv.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_anderson_coeff.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_aar.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, _dp, _dp,
                              C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, _dp]
        L.orc_poisson_solve.argtypes = [C.POINTER(OrcGrid), _dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_smearing.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_smearing.restype = C.c_double
        L.orc_fermi_level.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_double, C.c_double, C.c_double, C.c_int, _ip]
        L.orc_fermi_level.restype = C.c_double
        L.orc_density.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_double, C.c_double, _dp]
        L.orc_lda_pz.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_lda_pw.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def fd_weights(fdn: int = 6) -> np.ndarray:
    w = np.zeros(fdn + 1)
    assert lib().orc_fd_d2_weights(fdn, _p(w)) == 0
    return w


def mesh2chebdegree(h: float) -> int:
    return lib().orc_mesh2chebdegree(h)


class Oracle:
    """Oracle bound to one grid, Veff and projector set (mirrors a context)."""

    def __init__(self, grid, veff=None, nloc=None):
        self.grid = grid
        self.g = OrcGrid()
        st = lib().orc_grid_init(C.byref(self.g), grid.nx, grid.ny, grid.nz, grid.hx, grid.hy,
                                 grid.hz, grid.bcx, grid.bcy, grid.bcz, grid.order)
        if st != 0:
            raise ValueError(f"orc_grid_init failed ({st})")
        self.nd = grid.nd
        self.veff = None if veff is None else np.ascontiguousarray(veff, np.float64).ravel()
        self._nl = None
        self._nl_keep = None
        if nloc is not None and nloc.n_influence > 0:
            k = [np.ascontiguousarray(nloc.ndc, np.int32), np.ascontiguousarray(nloc.nproj, np.int32),
                 np.ascontiguousarray(nloc.pos_off, np.int64),
                 np.ascontiguousarray(nloc.grid_pos, np.int32),
                 np.ascontiguousarray(nloc.chi_off, np.int64), np.ascontiguousarray(nloc.chi, np.float64),
                 np.ascontiguousarray(nloc.alpha_off, np.int32),
                 np.ascontiguousarray(nloc.gamma, np.float64)]
            self._nl_keep = k
            self._nl = OrcNloc(nloc.n_influence, k[0].ctypes.data_as(_ip), k[1].ctypes.data_as(_ip),
                               k[2].ctypes.data_as(_lp), k[3].ctypes.data_as(_ip),
                               k[4].ctypes.data_as(_lp), _p(k[5]), k[6].ctypes.data_as(_ip),
                               int(nloc.n_alpha), _p(k[7]), float(nloc.dV))

    @property
    def nlp(self):
        return C.byref(self._nl) if self._nl is not None else None

    @property
    def coeffs(self):
        f = self.g.fdn + 1
        return (np.array(self.g.d2x[:f]), np.array(self.g.d2y[:f]), np.array(self.g.d2z[:f]))

    def max_eig_mhalf_lap(self) -> float:
        return lib().orc_max_eig_mhalf_lap(C.byref(self.g))

    @staticmethod
    def _block(x, nd):
        x = np.asfortranarray(np.asarray(x, np.float64))
        if x.ndim == 1:
            x = x.reshape(nd, 1, order="F")
        return x

    def lap_plus_diag(self, x, a, b, c, v=None):
        x = self._block(x, self.nd)
        y = np.zeros_like(x)
        vv = None if v is None else np.ascontiguousarray(v, np.float64).ravel()
        lib().orc_lap_plus_diag(C.byref(self.g), x.shape[1], a, b, c, _p(vv), _p(x), x.shape[0],
                                _p(y), y.shape[0])
        return y

    def lap_vec_mult(self, x, c=0.0):
        return self.lap_plus_diag(x, 1.0, 0.0, c)

    def vnl_vec_mult(self, x, Hx=None):
        x = self._block(x, self.nd)
        out = np.zeros_like(x) if Hx is None else np.asfortranarray(np.array(Hx, np.float64))
        if self._nl is not None:
            lib().orc_vnl_vec_mult(self.nlp, x.shape[1], _p(x), x.shape[0], _p(out), out.shape[0])
        return out

    def hamiltonian(self, x, c=0.0):
        x = self._block(x, self.nd)
        y = np.zeros_like(x)
        lib().orc_hamiltonian(C.byref(self.g), _p(self.veff), self.nlp, x.shape[1], c, _p(x),
                              x.shape[0], _p(y), y.shape[0])
        return y

    def chebyshev_filter(self, X, m, a, b, a0):
        X = np.asfortranarray(np.array(self._block(X, self.nd), copy=True))
        Y = np.zeros_like(X)
        lib().orc_chebyshev_filter(C.byref(self.g), _p(self.veff), self.nlp, _p(X), X.shape[0],
                                   _p(Y), Y.shape[0], X.shape[1], m, a, b, a0)
        return Y, X

    def lanczos(self, x0, tol_min, tol_max, maxit):
        x0 = np.ascontiguousarray(x0, np.float64).ravel()
        emin, emax = C.c_double(), C.c_double()
        it = lib().orc_lanczos(C.byref(self.g), _p(self.veff), self.nlp, _p(x0), tol_min, tol_max,
                               maxit, C.byref(emin), C.byref(emax))
        return emin.value, emax.value, it

    def project(self, Y):
        Y = self._block(Y, self.nd)
        n = Y.shape[1]
        HY = np.zeros_like(Y)
        Hp = np.zeros((n, n), order="F")
        Mp = np.zeros((n, n), order="F")
        lib().orc_project_hamiltonian(C.byref(self.g), _p(self.veff), self.nlp, _p(Y), Y.shape[0],
                                      _p(HY), HY.shape[0], n, _p(Hp), _p(Mp))
        return Hp, Mp, HY

    def rotate(self, Y, Q):
        Y = self._block(Y, self.nd)
        Q = np.asfortranarray(Q, np.float64)
        out = np.zeros_like(Y)
        lib().orc_subspace_rotation(self.nd, Y.shape[1], _p(Y), Y.shape[0], _p(Q), _p(out),
                                    out.shape[0])
        return out

    def poisson_solve(self, phi0, rhs, tol, maxit):
        phi = np.array(phi0, np.float64, copy=True).ravel()
        rhs = np.ascontiguousarray(rhs, np.float64).ravel()
        rr = C.c_double()
        it = lib().orc_poisson_solve(C.byref(self.g), _p(phi), _p(rhs), tol, maxit, C.byref(rr))
        return phi, it, rr.value

    def aar(self, x0, b, c=0.0, omega=0.6, beta=0.6, m=7, p=6, tol=1e-8, max_iter=1000):
        """AAR with the oracle's Poisson residual and Jacobi preconditioner,
        without the periodic mean shift.  Returns (x, iterations, relres)."""
        x = np.array(x0, np.float64, copy=True).ravel()
        b = np.ascontiguousarray(b, np.float64).ravel()
        rr = C.c_double()
        L = lib()
        res = C.cast(L.orc_poisson_residual, C.c_void_p)
        pre = C.cast(L.orc_jacobi_precond, C.c_void_p)
        it = L.orc_aar(C.cast(C.pointer(self.g), C.c_void_p), res, pre, c, self.nd, _p(x), _p(b),
                       omega, beta, m, p, tol, max_iter, C.byref(rr))
        return x, it, rr.value


def sygv(Hp, Mp):
    """Oracle generalized eigensolve (Cholesky + Jacobi).  Returns (lambda, Q)."""
    H = np.asfortranarray(np.array(Hp, np.float64, copy=True))
    M = np.asfortranarray(np.array(Mp, np.float64, copy=True))
    n = H.shape[0]
    lam, Q = np.zeros(n), np.zeros((n, n), order="F")
    st = lib().orc_sygv(n, _p(H), _p(M), _p(lam), _p(Q))
    if st != 0:
        raise np.linalg.LinAlgError(f"orc_sygv: Mp not positive definite (minor {st})")
    return lam, Q


def anderson_coeff(f, F):
    f = np.ascontiguousarray(f, np.float64)
    F = np.asfortranarray(F, np.float64)
    G = np.zeros(F.shape[1])
    lib().orc_anderson_coeff(F.shape[0], F.shape[1], _p(f), _p(F), _p(G))
    return G


def tridiag_minmax(d, e):
    d = np.ascontiguousarray(d, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    lo, hi = C.c_double(), C.c_double()
    lib().orc_tridiag_minmax(d.size, _p(d), _p(e), C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def fermi_level(lam, occfac, nelec, beta, smear_type, x1, x2, tol=1e-12, max_iter=100):
    lam = np.ascontiguousarray(lam, np.float64)
    st = C.c_int()
    ef = lib().orc_fermi_level(lam.size, _p(lam), occfac, nelec, beta, smear_type, x1, x2, tol,
                               max_iter, C.byref(st))
    return ef, st.value


def smearing(beta, lam, ef, smear_type):
    return np.array([lib().orc_smearing(beta, float(l), ef, smear_type) for l in np.ravel(lam)])


def density(psi, occ, occfac, dV):
    psi = np.asfortranarray(psi, np.float64)
    occ = np.ascontiguousarray(occ, np.float64)
    rho = np.zeros(psi.shape[0])
    lib().orc_density(psi.shape[0], psi.shape[1], _p(psi), psi.shape[0], _p(occ), occfac, dV,
                      _p(rho))
    return rho


def lda_pz(rho):
    rho = np.ascontiguousarray(rho, np.float64)
    exc, vxc = np.zeros_like(rho), np.zeros_like(rho)
    lib().orc_lda_pz(rho.size, _p(rho), _p(exc), _p(vxc))
    return exc, vxc


def lda_pw(rho):
    rho = np.ascontiguousarray(rho, np.float64)
    exc, vxc = np.zeros_like(rho), np.zeros_like(rho)
    lib().orc_lda_pw(rho.size, _p(rho), _p(exc), _p(vxc))
    return exc, vxc


def lda(rho, xc: int = 0):
    """xc 0: LDA_PZ, 1: LDA_PW."""
    return lda_pw(rho) if xc == 1 else lda_pz(rho)


# ------------------------------------------------------------ closed forms


def stencil_symbol(grid, coeffs):
    """Eigenvalues of the periodic FD Laplacian for every plane wave
    exp(i k.r): sum_axes (w0 + 2 sum_p w_p cos(k p h)) on the FFT frequency grid."""
    d2x, d2y, d2z = coeffs
    out = []
    for n, w in ((grid.nx, d2x), (grid.ny, d2y), (grid.nz, d2z)):
        th = 2 * np.pi * np.fft.fftfreq(n)
        s = np.full(n, w[0])
        for p in range(1, len(w)):
            s = s + 2 * w[p] * np.cos(p * th)
        out.append(s)
    sx, sy, sz = out
    return sz[:, None, None] + sy[None, :, None] + sx[None, None, :]


def fft_poisson(grid, coeffs, rhs):
    """Exact discrete solution of (Lap) phi = -rhs with zero mean (periodic)."""
    lam = stencil_symbol(grid, coeffs)
    R = np.fft.fftn(np.asarray(rhs).reshape(grid.nz, grid.ny, grid.nx))
    lam0 = lam.copy()
    lam0[0, 0, 0] = 1.0
    Phi = -R / lam0
    Phi[0, 0, 0] = 0.0
    return np.real(np.fft.ifftn(Phi)).ravel()


# ------------------------------------------------------------ reference build

_ref = None


def ref_lib():
    """The reference's own Lap_plus_diag_vec_mult_orth, or None if not built.

    Loaded with lazy binding: the library leaves the non-orthogonal / k-point
    helpers its file also defines unresolved, and they are never called."""
    global _ref
    if _ref is None and REF_LIB.exists():
        L = C.CDLL(str(REF_LIB), mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        L.ref_lap_plus_diag.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                        C.c_double, _dp, _dp, C.c_int, _dp, C.c_int]
        if hasattr(L, "ref_gradient"):
            L.ref_gradient.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       _dp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int]
        _ref = L
    return _ref


def ref_lap_plus_diag(grid, coeffs, x, a, b, c, v=None):
    L = ref_lib()
    if L is None:
        return None
    d2x, d2y, d2z = (np.ascontiguousarray(w, np.float64) for w in coeffs)
    x = np.asfortranarray(np.asarray(x, np.float64).reshape(grid.nd, -1))
    y = np.zeros_like(x)
    vv = None if v is None else np.ascontiguousarray(v, np.float64).ravel()
    L.ref_lap_plus_diag(grid.nx, grid.ny, grid.nz, grid.bcx, grid.bcy, grid.bcz, grid.order,
                        _p(d2x), _p(d2y), _p(d2z), x.shape[1], a, b, c, _p(vv), _p(x), x.shape[0],
                        _p(y), y.shape[0])
    return y


def ref_gradient(grid, x, axis):
    """The reference's own Gradient_vec_dir (c = 0) along axis (0 x, 1 y,
    2 z), or None where oracle/_ref is not built."""
    L = ref_lib()
    if L is None or not hasattr(L, "ref_gradient"):
        return None
    from .forces import d1_weights
    w = [np.ascontiguousarray(d1_weights(grid.order, h)) for h in (grid.hx, grid.hy, grid.hz)]
    x = np.asfortranarray(np.asarray(x, np.float64).reshape(grid.nd, -1))
    y = np.zeros_like(x)
    st = L.ref_gradient(grid.nx, grid.ny, grid.nz, grid.bcx, grid.bcy, grid.bcz, grid.order,
                        _p(w[0]), _p(w[1]), _p(w[2]), x.shape[1], axis, _p(x), x.shape[0], _p(y),
                        y.shape[0])
    if st != 0:
        raise RuntimeError("ref_gradient failed")
    return y
