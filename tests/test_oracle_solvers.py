"""Pin the oracle's filter, Lanczos, eigensolve, AAR, occupation, density and
LDA against closed forms, textbook library routines and brute force."""
import numpy as np
import pytest
import scipy.linalg as sla
from numpy.polynomial import chebyshev as C
from scipy import optimize, special

from oracle import oracle as O
from sparc_amd.workloads import Grid, poisson_workload, uniform_block
from tests.helpers import rel_err, small_case


@pytest.mark.parametrize("m", [1, 2, 5, 20])
def test_chebyshev_filter_closed_form_on_eigenvectors(m):
    """Veff constant, no projectors: plane waves are eigenvectors with
    eigenvalue lam = -sym/2 + V0, and the filter returns
    T_m((lam-c)/e) / T_m((a0-c)/e) x (Y) and the degree m-1 ratio (X)."""
    g = Grid(10, 9, 8, 0.45, 0.5, 0.55)
    V0 = -0.37
    o = O.Oracle(g, np.full(g.nd, V0))
    sym = O.stencil_symbol(g, o.coeffs)
    X, Y, Z = g.coords()
    a0, b = -2.0 + V0, float(-0.5 * sym.min()) + V0 + 0.5
    a = 0.5 * (a0 + b) - 3.0
    e, c = 0.5 * (b - a), 0.5 * (b + a)
    ks = [(0, 0, 0), (1, 0, 0), (2, 3, 1), (4, 4, 3)]
    cols, lams = [], []
    for kx, ky, kz in ks:
        k = 2 * np.pi * np.array([kx / (g.nx * g.hx), ky / (g.ny * g.hy), kz / (g.nz * g.hz)])
        cols.append(np.cos(k[0] * X + k[1] * Y + k[2] * Z))
        lams.append(-0.5 * sym[kz, ky, kx] + V0)
    Xin = np.stack(cols, axis=1)
    Yo, Xo = o.chebyshev_filter(Xin, m, a, b, a0)
    Tm = lambda deg, t: C.chebval(t, [0] * deg + [1])
    for j, lam in enumerate(lams):
        fy = Tm(m, (lam - c) / e) / Tm(m, (a0 - c) / e)
        fx = Tm(m - 1, (lam - c) / e) / Tm(m - 1, (a0 - c) / e) if m > 1 else 1.0
        assert np.max(np.abs(Yo[:, j] - fy * Xin[:, j])) < 1e-10 * max(1, abs(fy))
        assert np.max(np.abs(Xo[:, j] - fx * Xin[:, j])) < 1e-10 * max(1, abs(fx))


def test_lanczos_brackets_dense_spectrum():
    g, veff, nl, rng = small_case(n=(8, 7, 6), n_sites=2, rc=1.6, seed=7)
    o = O.Oracle(g, veff, nl)
    H = o.hamiltonian(np.eye(g.nd))
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))
    x0 = uniform_block(g.nd, 1, 3, 0.0, 1.0)[:, 0]
    emin, emax, it = o.lanczos(x0, 1e-10, 1e-10, 1000)
    assert abs(emin - ev[0]) < 1e-7 and abs(emax - ev[-1]) < 1e-7
    emin2, emax2, it2 = o.lanczos(x0, 1e-2, 1e-2, 1000)
    # Ritz values lie inside the spectrum
    assert ev[0] - 1e-12 <= emin2 and emax2 <= ev[-1] + 1e-12
    assert it2 <= it


def test_tridiag_extremes_vs_lapack():
    rng = np.random.default_rng(8)
    for n in (1, 2, 5, 40):
        d, e = rng.standard_normal(n), rng.standard_normal(max(n - 1, 0))
        ev = sla.eigvalsh_tridiagonal(d, e) if n > 1 else d
        lo, hi = O.tridiag_minmax(d, np.append(e, 0.0))
        assert abs(lo - ev[0]) < 1e-12 and abs(hi - ev[-1]) < 1e-12


def test_generalized_eigensolve_matches_lapack_dsygvd():
    rng = np.random.default_rng(9)
    n = 24
    A = rng.standard_normal((n, n)); Hp = A + A.T
    B = rng.standard_normal((n, n)); Mp = B @ B.T + n * np.eye(n)
    lam, Q = O.sygv(Hp, Mp)
    lref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.max(np.abs(lam - lref)) < 1e-11
    assert np.allclose(Q.T @ Mp @ Q, np.eye(n), atol=1e-11)
    assert np.allclose(Hp @ Q, Mp @ Q * lam, atol=1e-10)


def test_project_and_rotate_are_gemms():
    g, veff, nl, rng = small_case(n=(9, 8, 7), n_sites=2, rc=1.6, seed=10)
    o = O.Oracle(g, veff, nl)
    Y = rng.standard_normal((g.nd, 6))
    Hp, Mp, HY = o.project(Y)
    H = o.hamiltonian(np.eye(g.nd))
    assert rel_err(Mp, Y.T @ Y) < 1e-13
    assert rel_err(Hp, Y.T @ H @ Y) < 1e-12
    Q = rng.standard_normal((6, 6))
    assert rel_err(o.rotate(Y, Q), Y @ Q) < 1e-13


def test_anderson_coeff_is_min_norm_least_squares():
    """dgelsd(rcond=-1) on F'F: min-norm solution, also with zero history
    columns (the first Anderson step sees only p-1 of m columns filled)."""
    rng = np.random.default_rng(11)
    N, m = 200, 7
    F = rng.standard_normal((N, m))
    F[:, 5:] = 0.0
    f = rng.standard_normal(N)
    G = O.anderson_coeff(f, F)
    FtF, Ftf = F.T @ F, F.T @ f
    ref = np.linalg.lstsq(FtF, Ftf, rcond=None)[0]
    assert np.max(np.abs(G - ref)) < 1e-10
    assert np.all(G[5:] == 0.0)


@pytest.mark.parametrize("n", [12, 20])
def test_aar_poisson_matches_fft_solution(n):
    g, rhs = poisson_workload(n=n, h=0.4, n_charges=5, seed=n)
    o = O.Oracle(g)
    x, it, rr = o.aar(np.zeros(g.nd), rhs, tol=1e-10, max_iter=3000)
    assert rr <= 1e-10
    exact = O.fft_poisson(g, o.coeffs, rhs)
    assert rel_err(x - x.mean(), exact) < 1e-7
    phi, it2, rr2 = o.poisson_solve(np.zeros(g.nd), rhs, 1e-10, 3000)
    assert abs(phi.mean()) < 1e-12 and it2 == it


def test_aar_checks_residual_only_on_anderson_steps():
    g, rhs = poisson_workload(n=12, h=0.4, n_charges=3, seed=1)
    o = O.Oracle(g)
    for p in (1, 3, 6):
        _, it, _ = o.aar(np.zeros(g.nd), rhs, p=p, tol=1e-6, max_iter=2000)
        assert it % p == 0


@pytest.mark.parametrize("stype,beta", [(0, 100.0), (1, 7.349)])
def test_fermi_level_and_smearing(stype, beta):
    rng = np.random.default_rng(12)
    lam = np.sort(rng.uniform(-0.5, 0.6, 40))
    nelec = 37.0
    ef, st = O.fermi_level(lam, 2.0, nelec, beta, stype, lam[0] - 1, lam[-1] + 1)
    assert st == 0
    g = O.smearing(beta, lam, ef, stype)
    ref = special.expit(-beta * (lam - ef)) if stype == 0 else 0.5 * special.erfc(beta * (lam - ef))
    assert np.max(np.abs(g - ref)) < 1e-14
    assert abs(2.0 * g.sum() - nelec) < 1e-10
    f = lambda mu: 2.0 * (special.expit(-beta * (lam - mu)) if stype == 0 else
                          0.5 * special.erfc(beta * (lam - mu))).sum() - nelec
    assert abs(ef - optimize.brentq(f, lam[0] - 1, lam[-1] + 1, xtol=1e-14)) < 1e-9


def test_fermi_level_reports_unbracketable():
    lam = np.linspace(-1, 0, 4)
    ef, st = O.fermi_level(lam, 2.0, 20.0, 10.0, 1, -2.0, 1.0)  # 20 electrons in 4 states
    assert st == 1


def test_density_integrates_to_electron_count():
    rng = np.random.default_rng(13)
    nd, ns, dV = 300, 8, 0.07
    psi = np.linalg.qr(rng.standard_normal((nd, ns)))[0]
    occ = rng.uniform(0, 1, ns)
    rho = O.density(psi, occ, 2.0, dV)
    assert abs(rho.sum() * dV - 2.0 * occ.sum()) < 1e-12
    assert rel_err(rho, (psi ** 2 @ (2.0 * occ)) / dV) < 1e-14


def test_lda_pz_closed_forms_and_potential_derivative():
    rho = np.logspace(-5, 1, 60)
    exc, vxc = O.lda_pz(rho)
    ex = -0.75 * (3 / np.pi) ** (1 / 3) * np.cbrt(rho)
    rs = np.cbrt(3 / (4 * np.pi * rho))
    ec = np.where(rs < 1, 0.0311 * np.log(rs) - 0.048 + 0.002 * rs * np.log(rs) - 0.0116 * rs,
                  -0.1423 / (1 + 1.0529 * np.sqrt(rs) + 0.3334 * rs))
    assert np.max(np.abs(exc - (ex + ec))) < 1e-14
    # V_xc = d(rho * eps_xc)/d rho, checked by central differences
    h = 1e-6 * rho
    ep, _ = O.lda_pz(rho + h)
    em, _ = O.lda_pz(rho - h)
    dnum = ((rho + h) * ep - (rho - h) * em) / (2 * h)
    assert np.max(np.abs(dnum - vxc) / np.abs(vxc)) < 1e-6
