"""GPU parity of the Chebyshev filter, Rayleigh-Ritz pieces, Lanczos and a
full CheFSI iteration against the FP64 oracle (through the C ABI)."""
import numpy as np
import pytest
import scipy.linalg as sla

from tests.helpers import rel_err, small_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from sparc_amd import gpu
    gpu.load()
    return gpu


def _bounds(o, g, seed=11):
    from sparc_amd.workloads import uniform_block
    x0 = uniform_block(g.nd, 1, seed, 0.0, 1.0)[:, 0]
    emin, emax, _ = o.lanczos(x0, 1e-2, 1e-2, 1000)
    emax *= 1.01
    emin -= 0.1
    return emin, emax, 0.5 * (emin + emax)


@pytest.mark.parametrize("m", [1, 2, 7, 20])
def test_filter_fp64(gpu, oracle_mod, m):
    g, veff, nl, rng = small_case(n=(16, 15, 14), n_sites=3, seed=21)
    o = oracle_mod.Oracle(g, veff, nl)
    a0, b, a = _bounds(o, g)
    X = rng.uniform(-0.5, 0.5, (g.nd, 10))
    Yo, Xo = o.chebyshev_filter(X, m, a, b, a0)
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        Y, Xc = ctx.ChebyshevFiltering(X, m, a, b, a0, gpu.F64)
    assert rel_err(Y, Yo) < 1e-11
    assert rel_err(Xc, Xo) < 1e-11


def test_filter_fp32(gpu, oracle_mod):
    g, veff, nl, rng = small_case(n=(20, 19, 18), n_sites=4, seed=22)
    o = oracle_mod.Oracle(g, veff, nl)
    a0, b, a = _bounds(o, g)
    X = rng.uniform(-0.5, 0.5, (g.nd, 16))
    Yo, _ = o.chebyshev_filter(X, 20, a, b, a0)
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        Y, _ = ctx.ChebyshevFiltering(X, 20, a, b, a0, gpu.F32)
    # FP32 recurrence of degree 20: tolerance stated against the FP64 oracle
    assert rel_err(Y, Yo) < 1e-4


def test_filter_dev_storage_exchange(gpu, oracle_mod):
    """Even degree: the device filter swaps storage instead of copying."""
    g, veff, nl, rng = small_case(n=(12, 12, 12), n_sites=2, seed=23)
    o = oracle_mod.Oracle(g, veff, nl)
    a0, b, a = _bounds(o, g)
    X = rng.uniform(-0.5, 0.5, (g.nd, 5))
    Yo, Xo = o.chebyshev_filter(X, 6, a, b, a0)
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        bx = ctx.block(5).upload(X)
        by = ctx.block(5)
        ctx.chebyshev_filter_dev(bx, by, 6, a, b, a0)
        assert rel_err(by.download(), Yo) < 1e-11
        assert rel_err(bx.download(), Xo) < 1e-11


@pytest.mark.parametrize("dtype,tol", [(0, 1e-12), (1, 2e-5)])
def test_project_rotate(gpu, oracle_mod, dtype, tol):
    g, veff, nl, rng = small_case(n=(16, 16, 15), n_sites=3, seed=24)
    o = oracle_mod.Oracle(g, veff, nl)
    Y = rng.uniform(-0.5, 0.5, (g.nd, 12))
    Hpo, Mpo, HYo = o.project(Y)
    Q = np.linalg.qr(rng.standard_normal((12, 12)))[0]
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        Yin = Y.astype(np.float32).astype(np.float64) if dtype == 1 else Y
        Hp, Mp, HY = ctx.DP_Project_Hamiltonian(Yin, dtype)
        R = ctx.DP_Subspace_Rotation(Yin, Q, dtype)
    assert rel_err(Mp, Mpo) < tol
    assert rel_err(Hp, Hpo) < tol
    assert rel_err(HY, HYo) < max(tol, 1e-12)
    assert rel_err(R, o.rotate(Y, Q)) < max(tol, 1e-12)


def test_generalized_eigensolve(gpu, oracle_mod):
    rng = np.random.default_rng(25)
    n = 40
    A = rng.standard_normal((n, n)); Hp = A + A.T
    B = rng.standard_normal((n, n)); Mp = B @ B.T + n * np.eye(n)
    from sparc_amd.workloads import Grid
    with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
        lam, Q = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    lam_ref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.max(np.abs(lam - lam_ref)) < 1e-10
    assert np.allclose(Q.T @ Mp @ Q, np.eye(n), atol=1e-10)
    assert np.allclose(Hp @ Q, Mp @ Q * lam, atol=1e-9)
    lo, _ = oracle_mod.sygv(Hp, Mp)
    assert np.max(np.abs(lam - lo)) < 1e-10


def test_lanczos_matches_oracle(gpu, oracle_mod):
    from sparc_amd.workloads import uniform_block
    g, veff, nl, rng = small_case(n=(14, 14, 14), n_sites=3, seed=26)
    o = oracle_mod.Oracle(g, veff, nl)
    x0 = uniform_block(g.nd, 1, 5, 0.0, 1.0)[:, 0]
    emin_o, emax_o, it_o = o.lanczos(x0, 1e-2, 1e-2, 1000)
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        emin, emax, it = ctx.Lanczos(x0, 1e-2, 1e-2, 1000)
        e10 = ctx.Lanczos(x0, 0.0, 0.0, 10)
    assert it == it_o
    assert abs(emin - emin_o) < 1e-9 and abs(emax - emax_o) < 1e-9
    assert e10[2] == 10
    e10o = o.lanczos(x0, 0.0, 0.0, 10)
    assert abs(e10[0] - e10o[0]) < 1e-9 and abs(e10[1] - e10o[1]) < 1e-9


@pytest.mark.parametrize("dtype,tol", [(0, 1e-8), (1, 5e-5)])
def test_chefsi_converges_to_dense_spectrum(gpu, oracle_mod, dtype, tol):
    g, veff, nl, rng = small_case(n=(12, 12, 12), h=(0.5, 0.5, 0.5), n_sites=3, seed=27)
    o = oracle_mod.Oracle(g, veff, nl)
    H = np.diag(veff) + (-0.5) * o.lap_vec_mult(np.eye(g.nd)) + nl.dense(g.nd)
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))
    ns = 16
    X = rng.uniform(-0.5, 0.5, (g.nd, ns))
    a0, b, a = _bounds(o, g)
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(nl)
        bx = ctx.block(ns, dtype).upload(X)
        by = ctx.block(ns, dtype)
        lam = None
        for it in range(25):
            lam = ctx.chefsi_dev(bx, by, 20, a, b, a0)
            a0, a = lam[0], lam[-1] + 0.1
    nconv = ns - 4
    assert np.max(np.abs(lam[:nconv] - ev[:nconv])) < tol


def test_generalized_eigensolve_filtered_scale(gpu):
    """n = 480 with the spread of scales a Chebyshev-filtered block gives
    (Mp diagonal over eight decades, nearly diagonal Hp): eigenvalues against
    LAPACK dsygvd, Mp-orthonormal eigenvectors; and the not-positive-definite
    error path."""
    rng = np.random.default_rng(31)
    n = 480
    d = 10.0 ** rng.uniform(-4, 4, n)
    lam0 = np.sort(rng.uniform(-0.3, 0.5, n))
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-3 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(lam0) + (F + F.T) / 2)
    from sparc_amd.workloads import Grid
    with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
        lam, Q = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
        with pytest.raises(gpu.SparcGPUError):
            ctx.DP_Solve_Generalized_EigenProblem(Hp, -Mp)
    lam_ref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.max(np.abs(lam - lam_ref)) < 1e-10
    assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9


@pytest.mark.parametrize("n", [2, 24, 160, 180])
def test_generalized_eigensolve_lds_tridiagonal(gpu, monkeypatch, n):
    """Up to 180 states the reduction to tridiagonal form runs in one
    workgroup's LDS (dsytd2 'L'), then dstedc + dormtr: eigenvalues against
    LAPACK dsygvd and the rocsolver dsyevd path (SPARC_AMD_LDS_TRIDIAG=0) to
    1e-10, Mp-orthonormal eigenvectors that diagonalize Hp, on the
    filtered-block scale spread."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(70 + n)
    d = 10.0 ** rng.uniform(-4, 4, n)
    lam0 = np.sort(rng.uniform(-0.3, 0.5, n))
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-3 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(lam0) + (F + F.T) / 2)
    g = Grid(8, 8, 8, 0.5, 0.5, 0.5)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("SPARC_AMD_LDS_TRIDIAG", flag)
        with gpu.Context(g) as ctx:
            out[flag] = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    lam, Q = out["1"]
    lam_ref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.max(np.abs(lam - lam_ref)) < 1e-10
    assert np.max(np.abs(lam - out["0"][0])) < 1e-10
    assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9
    assert np.max(np.abs(Q.T @ Hp @ Q - np.diag(lam))) < 1e-9


@pytest.mark.parametrize("ns", [24, 160, 300])
def test_fp32_gram_accuracy(gpu, ns):
    """The FP32 projection (BF16-piece MFMA kernel below 256 states, rocBLAS
    above) against FP64 products of the same FP32 inputs: FP32-level error
    (5e-6 of max |Mp|), which the SCF at TOL_SCF = 1e-6 needs."""
    from sparc_amd.workloads import Grid
    g = Grid(20, 20, 20, 0.5, 0.5, 0.5)
    rng = np.random.default_rng(ns)
    Y = rng.uniform(-0.5, 0.5, (g.nd, ns)) * np.exp(rng.uniform(-3, 3, ns))
    Y32 = Y.astype(np.float32).astype(np.float64)
    veff = np.zeros(g.nd)
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        Hp, Mp, _ = ctx.DP_Project_Hamiltonian(Y, gpu.F32)
    ref = Y32.T @ Y32
    assert np.max(np.abs(Mp - ref)) < 5e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("ns", [300, 480])
def test_fp32_project_rotate_many_states(gpu, oracle_mod, ns):
    """At 256 states and more the FP32 Gram runs on the 128 x 128 BF16-piece
    kernel (Mp: upper tiles only, mirrored) and the rotation on the BF16-piece
    strip kernel.  Against FP64 products of the same FP32 inputs: Mp and Hp to
    5e-6 / 1e-5 of their largest element, HY and Psi = Y Q to 2e-5 relative."""
    from sparc_amd.workloads import Grid
    g = Grid(20, 20, 20, 0.5, 0.5, 0.5)
    rng = np.random.default_rng(ns + 1)
    Y = rng.uniform(-0.5, 0.5, (g.nd, ns)) * np.exp(rng.uniform(-3, 3, ns))
    Y32 = Y.astype(np.float32).astype(np.float64)
    veff = rng.uniform(-1.0, 0.0, g.nd)
    o = oracle_mod.Oracle(g, veff, None)
    Q = np.linalg.qr(rng.standard_normal((ns, ns)))[0]
    with gpu.Context(g) as ctx:
        ctx.set_veff(veff)
        Hp, Mp, HY = ctx.DP_Project_Hamiltonian(Y32, gpu.F32)
        R = ctx.DP_Subspace_Rotation(Y32, Q, gpu.F32)
    HYo = o.hamiltonian(Y32)
    Mref = Y32.T @ Y32
    Href = Y32.T @ HYo
    assert np.max(np.abs(Mp - Mref)) < 5e-6 * np.max(np.abs(Mref))
    assert np.max(np.abs(Hp - Href)) < 1e-5 * np.max(np.abs(Href))
    assert rel_err(HY, HYo) < 2e-5
    assert rel_err(R, Y32 @ Q) < 2e-5


def _separate_sites_case(seed):
    """Four projector sites whose 1.6-Bohr spheres do not overlap (one
    colour in the sphere layout), one of them across the cell face and split
    into two image entries that share alpha rows."""
    from sparc_amd.workloads import Grid, gaussian_projectors, gaussian_veff
    from tests.helpers import split_images
    g = Grid(20, 20, 20, 0.5, 0.5, 0.5)
    rng = np.random.default_rng(seed)
    veff = gaussian_veff(g, 4, rng)
    sites = np.array([[2.5, 2.5, 2.5], [7.5, 7.5, 2.5], [2.5, 7.5, 7.5], [0.2, 5.0, 7.0]])
    nl = gaussian_projectors(g, sites, 1.6, sigmas=(1.0, 0.7), gammas=(1.3, -0.6), lmax=1)
    return g, veff, split_images(nl, 3), rng


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("overlap", [False, True])
@pytest.mark.parametrize("mode", ["brick", "brick_fused", "sphere"])
def test_filter_nonlocal_chain(gpu, oracle_mod, monkeypatch, mode, overlap, dtype):
    """The filter forms each step's nonlocal partials inside the previous
    step's apply (sphere layout with one colour; brick layout with
    SPARC_AMD_NL_CHAIN_BRICK=1).  Against the unfused path
    (SPARC_AMD_NL_CHAIN=0): bitwise equal; against the oracle: the usual
    tolerance.  Separate spheres (one colour: the sphere layout fuses) and
    overlapping ones (the sphere layout applies colour by colour, unfused)."""
    from tests.helpers import split_images
    monkeypatch.setenv("SPARC_AMD_NL_MODE", "sphere" if mode == "sphere" else "brick")
    monkeypatch.setenv("SPARC_AMD_NL_CHAIN_BRICK", "1" if mode == "brick_fused" else "0")
    if overlap:
        g, veff, nl, rng = small_case(n=(16, 15, 14), n_sites=5, rc=2.6, seed=24, lmax=1)
        nl = split_images(nl, 1)
    else:
        g, veff, nl, rng = _separate_sites_case(25)
    o = oracle_mod.Oracle(g, veff, nl)
    a0, b, a = _bounds(o, g)
    ncol = 37
    X = rng.uniform(-0.5, 0.5, (g.nd, ncol))
    Yo, _ = o.chebyshev_filter(X, 9, a, b, a0)
    out = {}
    for chain in ("1", "0"):
        monkeypatch.setenv("SPARC_AMD_NL_CHAIN", chain)
        with gpu.Context(g) as ctx:
            ctx.set_veff(veff)
            ctx.set_nonlocal(nl)
            bx = ctx.block(ncol, dtype).upload(X)
            by = ctx.block(ncol, dtype)
            ctx.chebyshev_filter_dev(bx, by, 9, a, b, a0)  # first call times the stencil tiles
            bx.upload(X)
            ctx.chebyshev_filter_dev(bx, by, 9, a, b, a0)
            out[chain] = by.download()
    assert np.array_equal(out["1"], out["0"])
    assert rel_err(out["1"], Yo) < (1e-11 if dtype == 0 else 1e-4)


@pytest.mark.parametrize("n,groups,threads", [(181, 32, 256), (300, 7, 256), (480, 32, 256),
                                              (480, 64, 512), (1000, 16, 256)])
def test_generalized_eigensolve_cooperative_tridiagonal(gpu, monkeypatch, n, groups, threads):
    """Above 180 states the reduction to tridiagonal form runs in one
    cooperative launch, the rows spread over the workgroups' LDS (dsytd2 'L'
    with one grid barrier per column), then dstedc + dormtr: eigenvalues
    against LAPACK dsygvd and rocsolver's dsyevd path
    (SPARC_AMD_COOP_TRIDIAG=0) to 1e-10, Mp-orthonormal eigenvectors that
    diagonalize Hp, on the filtered-block scale spread; group counts that do
    not divide n, and n = 1000 (more groups than asked, for the LDS)."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(90 + n + groups)
    d = 10.0 ** rng.uniform(-4, 4, n)
    lam0 = np.sort(rng.uniform(-0.3, 0.5, n))
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-3 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(lam0) + (F + F.T) / 2)
    g = Grid(8, 8, 8, 0.5, 0.5, 0.5)
    out = {}
    monkeypatch.setenv("SPARC_AMD_COOP_THREADS", str(threads))
    for flag in (str(groups), "0"):
        monkeypatch.setenv("SPARC_AMD_COOP_TRIDIAG", flag)
        with gpu.Context(g) as ctx:
            out[flag] = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    lam, Q = out[str(groups)]
    lam_ref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.max(np.abs(lam - lam_ref)) < 1e-10
    assert np.max(np.abs(lam - out["0"][0])) < 1e-10
    assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9
    assert np.max(np.abs(Q.T @ Hp @ Q - np.diag(lam))) < 1e-9


def test_generalized_eigensolve_cooperative_degenerate(gpu):
    """Exactly degenerate eigenvalues (a cubic crystal's threefold states) and
    a zero column below the diagonal (the Householder step with tau = 0):
    n = 240 built as V diag(lam) V' with triplets, and a block-diagonal
    matrix whose first column is e_0."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(44)
    n = 240
    lam0 = np.repeat(np.sort(rng.uniform(-0.4, 0.6, n // 3)), 3)
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Hp = (V * lam0) @ V.T
    Hp = 0.5 * (Hp + Hp.T)
    Hb = np.zeros((n, n))
    Hb[0, 0] = 0.7
    B = rng.standard_normal((n - 1, n - 1))
    Hb[1:, 1:] = B + B.T
    Mp = np.eye(n)
    with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
        for H in (Hp, Hb):
            lam, Q = ctx.DP_Solve_Generalized_EigenProblem(H, Mp)
            assert np.max(np.abs(lam - np.linalg.eigvalsh(H))) < 1e-10
            assert np.max(np.abs(Q.T @ Q - np.eye(n))) < 1e-10
            assert np.max(np.abs(Q.T @ H @ Q - np.diag(lam))) < 1e-9


@pytest.mark.parametrize("n", [24, 181, 480])
def test_generalized_eigensolve_blocked_cholesky(gpu, monkeypatch, n):
    """The overlap factor from the blocked Cholesky kernels (panels of 32
    columns; n = 181 leaves a partial panel) with C = L^-1 Hp L^-T and
    Q = L^-T Z as products, against rocSOLVER's dpotrf + triangular solves
    (SPARC_AMD_CHOL=rocsolver) and LAPACK dsygvd: eigenvalues to 1e-10,
    Mp-orthonormal eigenvectors; a matrix that is not positive definite in
    its last panel raises, with the column of the failing pivot."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(120 + n)
    d = 10.0 ** rng.uniform(-4, 4, n)
    lam0 = np.sort(rng.uniform(-0.3, 0.5, n))
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-3 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(lam0) + (F + F.T) / 2)
    g = Grid(8, 8, 8, 0.5, 0.5, 0.5)
    out = {}
    for flag in ("own", "rocsolver"):
        monkeypatch.setenv("SPARC_AMD_CHOL", flag)
        with gpu.Context(g) as ctx:
            out[flag] = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    lam, Q = out["own"]
    lam_ref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.max(np.abs(lam - lam_ref)) < 1e-10
    assert np.max(np.abs(lam - out["rocsolver"][0])) < 1e-10
    assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9
    assert np.max(np.abs(Q.T @ Hp @ Q - np.diag(lam))) < 1e-9
    bad = np.eye(n)
    bad[n - 1, n - 1] = -1.0
    monkeypatch.setenv("SPARC_AMD_CHOL", "own")
    with gpu.Context(g) as ctx:
        with pytest.raises(gpu.SparcGPUError, match=f"info = {n}"):
            ctx.DP_Solve_Generalized_EigenProblem(np.eye(n), bad)
        lam2, _ = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)   # usable afterwards
    assert np.max(np.abs(lam2 - lam_ref)) < 1e-10


@pytest.mark.parametrize("n", [24, 181, 480, 1000])
def test_generalized_eigensolve_back_transformation(gpu, monkeypatch, n):
    """The back-transformation Q1 Z by one wave per four columns of Z (the
    reflectors applied from the last to the first in registers) against
    rocSOLVER's dormtr (SPARC_AMD_ORMTR=rocsolver): the same eigenvalues and
    eigenvectors to rounding (same tridiagonal matrix and dstedc), and both
    Mp-orthonormal and diagonalizing Hp; n = 24 takes the one-workgroup
    reduction, the others the cooperative one."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(200 + n)
    d = 10.0 ** rng.uniform(-3, 3, n)
    lam0 = np.sort(rng.uniform(-0.3, 0.5, n))
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-2 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(lam0) + (F + F.T) / 2)
    g = Grid(8, 8, 8, 0.5, 0.5, 0.5)
    out = {}
    for flag in ("own", "rocsolver"):
        monkeypatch.setenv("SPARC_AMD_ORMTR", flag)
        with gpu.Context(g) as ctx:
            out[flag] = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    (lam, Q), (lam_r, Q_r) = out["own"], out["rocsolver"]
    assert np.array_equal(lam, lam_r)
    scale = np.max(np.abs(Q_r))
    assert np.max(np.abs(Q - Q_r)) < 1e-11 * scale
    assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9
    assert np.max(np.abs(Q.T @ Hp @ Q - np.diag(lam))) < 1e-9


@pytest.mark.parametrize("n", [1200, 1600])
def test_generalized_eigensolve_large(gpu, n):
    """n = 1200: the cooperative reduction with more workgroups than asked
    (its rows need about 110 workgroups' LDS); n = 1600: past what 128
    workgroups' LDS holds, the reduction falls back to rocSOLVER's dsyevd.
    Eigenvalues against LAPACK dsygvd, Mp-orthonormal eigenvectors."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(300 + n)
    d = 10.0 ** rng.uniform(-2, 2, n)
    lam0 = np.sort(rng.uniform(-0.3, 0.5, n))
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-3 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(lam0) + (F + F.T) / 2)
    with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
        lam, Q = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    lam_ref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.max(np.abs(lam - lam_ref)) < 1e-10
    assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9
    assert np.max(np.abs(Q.T @ Hp @ Q - np.diag(lam))) < 1e-9


@pytest.mark.parametrize("kind", ["random", "clusters", "graded", "multiple"])
@pytest.mark.parametrize("n", [2, 24, 181, 480, 1000])
def test_generalized_eigensolve_divide_and_conquer(gpu, monkeypatch, n, kind):
    """The tridiagonal eigenproblem by the library's own divide and conquer
    (1 x 1 leaves, dlaed2 deflation, the secular equation by the middle-way
    iteration, Gu-Eisenstat vectors) against rocSOLVER's dstedc
    (SPARC_AMD_STEDC=rocsolver) and LAPACK dsygvd: eigenvalues to 1e-10 of
    the spectrum's scale, Mp-orthonormal eigenvectors that diagonalize Hp.
    Spectra: random; sixfold clusters (C = V diag V' exactly degenerate,
    deflation by rotation chains); graded over 8 decades; one eigenvalue of
    multiplicity n (Hp = 2 Mp: every coupling deflates)."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(500 + n + len(kind))
    d = 10.0 ** rng.uniform(-3, 3, n)
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    L = np.linalg.cholesky(Mp)
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if kind == "random":
        lam0 = rng.uniform(-0.3, 0.5, n)
    elif kind == "clusters":
        lam0 = np.repeat(rng.uniform(-0.3, 0.5, n // 6 + 1), 6)[:n]
    elif kind == "graded":
        lam0 = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-8, 0, n)
    else:
        lam0 = np.full(n, 2.0)
    C = (V * lam0) @ V.T
    Hp = L @ (0.5 * (C + C.T)) @ L.T
    Hp = 0.5 * (Hp + Hp.T)
    g = Grid(8, 8, 8, 0.5, 0.5, 0.5)
    out = {}
    for flag in ("own", "rocsolver"):
        monkeypatch.setenv("SPARC_AMD_STEDC", flag)
        with gpu.Context(g) as ctx:
            out[flag] = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    lam, Q = out["own"]
    scale = np.max(np.abs(lam0))
    lam_ref = sla.eigh(Hp, Mp, eigvals_only=True, driver="gvd")
    assert np.all(np.diff(lam) >= 0)
    assert np.max(np.abs(lam - lam_ref)) < 1e-10 * scale
    assert np.max(np.abs(lam - out["rocsolver"][0])) < 1e-10 * scale
    assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9
    assert np.max(np.abs(Q.T @ Hp @ Q - np.diag(lam))) < 1e-9 * scale


@pytest.mark.parametrize("ncol", [160, 256, 300])
def test_projection_pair_kernel_bitwise(gpu, monkeypatch, ncol):
    """Mp = Y'Y and Hp = Y'HY of an FP32 block from the one-launch pair kernel
    (default from 128 states) against the two single-product launches
    (SPARC_AMD_GRAM3_PAIR=0): bitwise equal from 256 states, where those are
    the same 128 x 128 kernel (300 columns leave partial tiles); at 160 states
    (the 64 x 64 kernel there) both FP32-accurate.  Mp against the FP64
    product of the FP32 values to 1e-5."""
    from sparc_amd.workloads import Grid
    g = Grid(24, 24, 24, 0.5, 0.5, 0.5)
    rng = np.random.default_rng(810 + ncol)
    Y = rng.standard_normal((g.nd, ncol))
    veff = rng.uniform(-1.0, 0.0, g.nd)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("SPARC_AMD_GRAM3_PAIR", flag)
        with gpu.Context(g) as ctx:
            ctx.set_veff(veff)
            by = ctx.block(ncol, gpu.F32).upload(Y)
            bh = ctx.block(ncol, gpu.F32)
            out[flag] = ctx.project_dev(by, bh)
    if ncol >= 256:
        assert np.array_equal(out["1"][0], out["0"][0]) and np.array_equal(out["1"][1], out["0"][1])
    else:
        for k in (0, 1):
            assert np.max(np.abs(out["1"][k] - out["0"][k])) < 1e-5 * np.max(np.abs(out["0"][k]))
    Yf = Y.astype(np.float32).astype(np.float64)
    Mref = Yf.T @ Yf
    assert np.max(np.abs(out["1"][1] - Mref)) < 1e-5 * np.max(np.abs(Mref))


def test_generalized_eigensolve_refuses_non_finite(gpu):
    """A NaN or infinity in Hp or Mp is refused on the host before any upload
    (LAPACK's dsygvd would report it through info), with the matrix and the
    entry named; the context stays usable."""
    from sparc_amd.workloads import Grid
    n = 32
    rng = np.random.default_rng(900)
    A = rng.standard_normal((n, n))
    Hp = A + A.T
    Mp = np.eye(n) + 0.01 * (A @ A.T)
    with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
        for bad_h in (True, False):
            H, M = Hp.copy(), Mp.copy()
            (H if bad_h else M)[3, 5] = np.nan if bad_h else np.inf
            with pytest.raises(gpu.SparcGPUError, match="non-finite entry in " + ("Hp" if bad_h else "Mp")):
                ctx.DP_Solve_Generalized_EigenProblem(H, M)
        lam, _ = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
    assert np.max(np.abs(lam - sla.eigh(Hp, Mp, eigvals_only=True))) < 1e-10 * np.max(np.abs(lam))


def _tridiagonal_case(kind, n, rng):
    """Symmetric tridiagonal test matrices whose spectra the mathematics fixes
    (Clement, 1-D Laplacian, diagonal) or that are hard for divide and
    conquer (Wilkinson: pairs of eigenvalues equal to many digits; glued
    Wilkinson: 21-point Wilkinson blocks coupled by 1e-12)."""
    if kind == "clement":        # eigenvalues -(n-1), -(n-3), ..., n-3, n-1
        d = np.zeros(n)
        e = np.sqrt(np.arange(1, n) * (n - np.arange(1, n)))
        exact = np.arange(-(n - 1), n, 2, dtype=float)
    elif kind == "laplacian":    # eigenvalues 2 - 2 cos(k pi / (n + 1))
        d = np.full(n, 2.0)
        e = np.full(n - 1, -1.0)
        exact = np.sort(2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
    elif kind == "diagonal":     # every coupling zero: the sorted diagonal
        d = rng.uniform(-1.0, 1.0, n)
        e = np.zeros(n - 1)
        exact = np.sort(d)
    elif kind == "wilkinson":
        m = (n - 1) // 2
        d = np.abs(np.arange(n) - m).astype(float)
        e = np.ones(n - 1)
        exact = None
    else:                        # glued Wilkinson W21 blocks
        w = np.abs(np.arange(21) - 10).astype(float)
        nb = n // 21
        d = np.concatenate([np.tile(w, nb), np.zeros(n - 21 * nb)])
        e = np.ones(n - 1)
        e[20:21 * nb - 1:21] = 1e-12
        exact = None
    return d, e, exact


@pytest.mark.parametrize("kind", ["clement", "laplacian", "diagonal", "wilkinson", "glued"])
@pytest.mark.parametrize("n", [150, 480])
def test_generalized_eigensolve_tridiagonal_matrices(gpu, kind, n):
    """With Mp = I and a tridiagonal Hp the divide and conquer receives that
    tridiagonal matrix (the reduction leaves it as it is, up to signs):
    eigenvalues against the closed forms (Clement's integers, the Laplacian's
    cosines, the sorted diagonal) or LAPACK via scipy (Wilkinson and glued
    Wilkinson) to 1e-10 of the spectrum's scale, orthonormal eigenvectors
    that diagonalize the matrix; n = 150 takes the one-workgroup reduction,
    480 the cooperative one."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(700 + n)
    d, e, exact = _tridiagonal_case(kind, n, rng)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    if exact is None:
        exact = sla.eigh_tridiagonal(d, e, eigvals_only=True)
    with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
        lam, Q = ctx.DP_Solve_Generalized_EigenProblem(T, np.eye(n))
    scale = np.max(np.abs(exact))
    assert np.all(np.diff(lam) >= 0)
    assert np.max(np.abs(lam - exact)) < 1e-10 * scale
    assert np.max(np.abs(Q.T @ Q - np.eye(n))) < 1e-9
    assert np.max(np.abs(Q.T @ T @ Q - np.diag(lam))) < 1e-9 * scale


@pytest.mark.parametrize("leaf", ["1", "16", "32"])
def test_generalized_eigensolve_divide_and_conquer_leaf(gpu, monkeypatch, leaf):
    """The lowest merge levels inside one wave per leaf block (default 8
    poles; SPARC_AMD_TDC_LEAF=1 runs every level as separate launches, 16 and
    32 larger leaves): the same eigenvalues as the default to 1e-12 and
    Mp-orthonormal eigenvectors that diagonalize Hp, for a random and a
    sixfold-clustered spectrum at n = 181 (partial last leaf)."""
    from sparc_amd.workloads import Grid
    n = 181
    rng = np.random.default_rng(610)
    d = 10.0 ** rng.uniform(-3, 3, n)
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    L = np.linalg.cholesky(Mp)
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    g = Grid(8, 8, 8, 0.5, 0.5, 0.5)
    for lam0 in (rng.uniform(-0.3, 0.5, n), np.repeat(rng.uniform(-0.3, 0.5, n // 6 + 1), 6)[:n]):
        C = (V * lam0) @ V.T
        Hp = L @ (0.5 * (C + C.T)) @ L.T
        Hp = 0.5 * (Hp + Hp.T)
        out = {}
        for flag in ("8", leaf):
            monkeypatch.setenv("SPARC_AMD_TDC_LEAF", flag)
            with gpu.Context(g) as ctx:
                out[flag] = ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
        lam, Q = out[leaf]
        assert np.max(np.abs(lam - out["8"][0])) < 1e-12
        assert np.max(np.abs(Q.T @ Mp @ Q - np.eye(n))) < 1e-9
        assert np.max(np.abs(Q.T @ Hp @ Q - np.diag(lam))) < 1e-9


@pytest.mark.parametrize("n", [480, 1100])
def test_generalized_eigensolve_repeatable(gpu, n):
    """Four solves of the same pair in one context give bitwise the same
    eigenvalues and eigenvectors: every sum of the Cholesky panels, the
    cooperative reduction and the back-transformation runs in a fixed order,
    so a difference would show a race (as between a panel's diagonal-block
    write and the row blocks still reading it)."""
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(400 + n)
    d = 10.0 ** rng.uniform(-3, 3, n)
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-2 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(np.sort(rng.uniform(-0.3, 0.5, n))) + (F + F.T) / 2)
    with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
        runs = [ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp) for _ in range(4)]
    for lam, Q in runs[1:]:
        assert np.array_equal(lam, runs[0][0])
        assert np.array_equal(Q, runs[0][1])
