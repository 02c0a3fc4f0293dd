"""GPU parity at the sizes of the five benchmark workloads.

8-atom Si (psp8, LDA-PZ), 26^3, 24 states, degree 20, TOL_SCF 1e-6: the
    FP64 device SCF against the FP64 oracle SCF (eigenvalues <= 1e-6 Ha,
    total energy <= 1e-5 Ha/atom, SCF count within 3 as the reference's test
    harness allows); the mixed run <= 1e-5 Ha/atom of the oracle.
Operator filter, 96^3 x 512, degree 24, 64 sites x 8 projectors (rc 6):
    the 10-step Lanczos bounds against the oracle's (1e-9 relative), then the
    full 512-column device filter in FP64 and FP32 against the FP64 oracle
    filter of 32 of its columns (every 16th): FP64 1e-11, FP32 1e-4 relative
    to max |Y| (both outputs, Y_m and the clobbered X = Y_{m-1}).
64-atom Si, 52^3, 160 states: one CheFSI pass (filter, projection, FP64
    eigensolve, rotation) from the seeded orbitals with the oracle's bounds
    and potential, eigenvalues against the oracle's pass (FP64 1e-8 Ha, FP32
    5e-5 Ha); then the FP64 and mixed SCF to convergence, mixed within
    1e-5 Ha/atom and 1e-5 Ha (eigenvalues) of FP64.
128^3 Poisson, 216 Gaussian charges, rel. tol 1e-8: Jacobi-AAR and
    U-Net-AAR against the exact FFT solution of the discrete problem
    (oracle.fft_poisson, 1e-5 relative), both at the residual bound, the
    U-Net run in fewer iterations.
216-atom Si/Al (10 % Al, Fermi-Dirac 0.01 Ha), 480 states: FP64, mixed
    and mixed + learned Poisson preconditioner SCFs converge; the mixed runs
    within 1e-5 Ha/atom of FP64; atomic forces of the mixed run within
    1e-6 Ha/Bohr of FP64 and of the learned run within 3e-4 (its Poisson
    solves stop at other iterates within TOL_POISSON); with every Poisson
    solve converged to 1e-10, learned and Jacobi forces agree to 1e-6.

The U-Net used by the Poisson and 216-atom tests is trained on the device,
as in the benchmark, on (residual, error) pairs of Jacobi-AAR Poisson solves
of the 8-atom and 64-atom cells' densities; the 216-atom cell runs on its
78^3 grid, the network on 80^3 through the interpolations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _step(msg):
    print(f"\n[configs] {msg}", flush=True)


# ----------------------------------------------------------------- si8

def test_si8_fp64_and_mixed_match_oracle_scf():
    from oracle import oracle as orc
    from oracle.scf import OracleSCF
    from sparc_amd import gpu
    from sparc_amd.systems import scf_system
    from sparc_amd.workloads import uniform_block
    orc.set_threads(16)
    st, p = scf_system("si8")
    p = dict(p, tol_scf=1e-6)
    assert (st.grid.nx, p["nstates"], p["cheb_degree"]) == (26, 24, 20)
    nd = st.grid.nd
    psi0 = uniform_block(nd, p["nstates"], p["psi_seed"], -0.5, 0.5)
    x0 = uniform_block(nd, 1, p["lanczos_seed"], 0.0, 1.0)
    _step("si8 oracle SCF")
    ref = OracleSCF(st.grid, st.nloc, st.b, st.rho_at, p, psi0, x0, rho_core=st.rho_core).run()
    out = {}
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        for dt in (gpu.F64, gpu.F32):
            _step(f"si8 device SCF dtype {dt}")
            out[dt] = ctx.scf_run(st.b, st.rho_at, dict(p, dtype=dt), rho_core=st.rho_core)
    r = out[gpu.F64]
    de = abs(r["Etot"] - ref["Etot"]) / st.n_atom
    dl = np.max(np.abs(r["lambda_"] - ref["lambda_"]))
    print(f"si8 FP64: {r['scf_iters']} iters (oracle {ref['scf_iters']}), dE/atom {de:.2e}, "
          f"max dlambda {dl:.2e}")
    assert r["converged"] == 1 and ref["converged"]
    assert abs(r["scf_iters"] - ref["scf_iters"]) <= 3
    assert de < 1e-5
    assert dl < 1e-6
    rm = out[gpu.F32]
    dem = abs(rm["Etot"] - ref["Etot"]) / st.n_atom
    print(f"si8 mixed: {rm['scf_iters']} iters, dE/atom vs oracle {dem:.2e}")
    assert rm["converged"] == 1
    assert dem < 1e-5


# ------------------------------------------------------------ filter96

def test_operator_filter_96cubed_512_degree24():
    from oracle import oracle as orc
    from sparc_amd import gpu
    from sparc_amd.workloads import operator_filter_workload, uniform_block
    orc.set_threads(16)
    wl = operator_filter_workload()
    g, m = wl.grid, wl.degree
    assert (g.nx, wl.ncol, m, wl.nloc.n_influence) == (96, 512, 24, 64)
    o = orc.Oracle(g, wl.veff, wl.nloc)
    x0 = uniform_block(g.nd, 1, wl.lanczos_seed, 0.0, 1.0)[:, 0]
    _step("filter96 oracle Lanczos")
    emin_o, emax_o, _ = o.lanczos(x0, 0.0, 0.0, 10)
    cols = np.arange(0, wl.ncol, 16)
    res = {}
    with gpu.Context(g) as ctx:
        ctx.set_veff(wl.veff)
        ctx.set_nonlocal(wl.nloc)
        emin, emax, _ = ctx.Lanczos(x0, 0.0, 0.0, 10)
        assert abs(emin - emin_o) <= 1e-9 * abs(emin_o)
        assert abs(emax - emax_o) <= 1e-9 * abs(emax_o)
        b, a0 = emax * 1.01, emin - 0.1
        a = 0.5 * (a0 + b)
        for dt in (gpu.F64, gpu.F32):
            _step(f"filter96 device filter dtype {dt}")
            X = ctx.block(wl.ncol, dt).fill_uniform(wl.psi_seed, -0.5, 0.5)
            Y = ctx.block(wl.ncol, dt)
            ctx.chebyshev_filter_dev(X, Y, m, a, b, a0)
            yd, xd = Y.download(), X.download()
            res[dt] = (yd[:, cols].copy(), xd[:, cols].copy())
            del yd, xd
            X.free()
            Y.free()
    X0 = uniform_block(g.nd, wl.ncol, wl.psi_seed, -0.5, 0.5)[:, cols].copy(order="F")
    _step(f"filter96 oracle filter of {cols.size} columns")
    Yo, Xo = o.chebyshev_filter(X0, m, a, b, a0)
    sy, sx = np.abs(Yo).max(), np.abs(Xo).max()
    for dt, tol in ((gpu.F64, 1e-11), (gpu.F32, 1e-4)):
        ey = np.abs(res[dt][0] - Yo).max() / sy
        ex = np.abs(res[dt][1] - Xo).max() / sx
        print(f"filter96 dtype {dt}: Y rel err {ey:.2e}, X rel err {ex:.2e} (bar {tol:g})")
        assert ey < tol and ex < tol


# ---------------------------------------------------------------- si64

@pytest.fixture(scope="module")
def si64():
    from sparc_amd.systems import scf_system
    return scf_system("si64")


def test_si64_chefsi_pass_matches_oracle(si64):
    from oracle import oracle as orc
    from oracle.scf import OracleSCF
    from sparc_amd import gpu
    from sparc_amd.workloads import uniform_block
    orc.set_threads(16)
    st, p = si64
    nd, ns, m = st.grid.nd, p["nstates"], p["cheb_degree"]
    assert (st.grid.nx, ns) == (52, 160)
    psi0 = uniform_block(nd, ns, p["psi_seed"], -0.5, 0.5)
    x0 = uniform_block(nd, 1, p["lanczos_seed"], 0.0, 1.0)
    _step("si64 oracle potential and CheFSI pass")
    s = OracleSCF(st.grid, st.nloc, st.b, st.rho_at, p, psi0, x0, rho_core=st.rho_core)
    veff = s.o.veff.copy()
    s.chefsi(0)
    a0, b = s.eigmin, s.eigmax
    a = 0.5 * (a0 + b)
    lam_o = s.lam.copy()
    lam = {}
    with gpu.Context(st.grid) as ctx:
        ctx.set_veff(veff)
        ctx.set_nonlocal(st.nloc)
        for dt in (gpu.F64, gpu.F32):
            _step(f"si64 device CheFSI pass dtype {dt}")
            X = ctx.block(ns, dt).fill_uniform(p["psi_seed"], -0.5, 0.5)
            Y = ctx.block(ns, dt)
            lam[dt] = ctx.chefsi_dev(X, Y, m, a, b, a0)
            X.free()
            Y.free()
    for dt, tol in ((gpu.F64, 1e-8), (gpu.F32, 5e-5)):
        d = np.max(np.abs(lam[dt] - lam_o))
        print(f"si64 pass dtype {dt}: max |dlambda| {d:.2e} Ha (bar {tol:g})")
        assert d < tol


def test_si64_mixed_scf_matches_fp64(si64):
    from sparc_amd import gpu
    st, p = si64
    out = {}
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        for dt in (gpu.F64, gpu.F32):
            _step(f"si64 device SCF dtype {dt}")
            out[dt] = ctx.scf_run(st.b, st.rho_at, dict(p, dtype=dt), rho_core=st.rho_core)
    f, mx = out[gpu.F64], out[gpu.F32]
    de = abs(mx["Etot"] - f["Etot"]) / st.n_atom
    dl = np.max(np.abs(mx["lambda_"] - f["lambda_"]))
    print(f"si64 SCF: FP64 {f['scf_iters']} iters {f['mean_iter_ms']:.2f} ms/iter, mixed "
          f"{mx['scf_iters']} iters {mx['mean_iter_ms']:.2f} ms/iter, dE/atom {de:.2e}, "
          f"max dlambda {dl:.2e}")
    assert f["converged"] == 1 and mx["converged"] == 1
    assert de < 1e-5
    assert dl < 1e-5


# ------------------------------------------ trained U-Net (Poisson, 216-atom SCF)

@pytest.fixture(scope="module")
def unet_weights():
    """The network trained as the benchmark trains it: pairs of the 8-atom
    (26^3, run on 28^3 by interpolation) and 64-atom (52^3) cells' Poisson
    solves, the earliest weights with the fewest AAR iterations summed over
    held-out 8- and 64-atom problems kept."""
    from sparc_amd.learned import train_poisson_unet
    _step("U-Net training on si8 + si64 pairs")
    w, info = train_poisson_unet(("si8", "si64"), steps=400)
    print(f"pairs {info['pairs']}, validation {info['validation_aar_iterations']}, kept after "
          f"{info['kept_after_steps']} steps")
    assert info["loss_last"] < 0.2 * info["loss_first"]
    assert info["kept_after_steps"] > 0   # some training beat the zero network (Jacobi)
    return w


# ---------------------------------------------------------- poisson128

def test_poisson_128cubed_jacobi_and_unet_match_fft(unet_weights):
    from oracle import oracle as orc
    from sparc_amd import gpu
    from sparc_amd.workloads import poisson_workload
    g, rhs = poisson_workload()
    assert g.nx == 128
    tol = 1e-8
    with gpu.Context(g) as ctx:
        net = ctx.unet(seed=1)
        net.set_weights(unet_weights)
        x0 = np.zeros(g.nd)
        _step("poisson128 Jacobi-AAR")
        xj, itj, rrj, okj = ctx.AAR(x0, rhs, tol=tol, max_iter=10000)
        _step("poisson128 U-Net-AAR")
        xu, itu, rru, stu = ctx.AAR_unet(net, x0, rhs, tol=tol, max_iter=10000)
        # residuals of both solutions on the device stencil (r = b + Lap x)
        resj = np.linalg.norm(rhs + ctx.Lap_vec_mult(xj).ravel())
        resu = np.linalg.norm(rhs + ctx.Lap_vec_mult(xu).ravel())
        net.close()
    exact = orc.fft_poisson(g, orc.Oracle(g).coeffs, rhs)
    nb = np.linalg.norm(rhs)
    for name, x, it, res in (("Jacobi", xj, itj, resj), ("U-Net", xu, itu, resu)):
        x = x - x.mean()
        err = np.linalg.norm(x - exact) / np.linalg.norm(exact)
        print(f"poisson128 {name}: {it} iterations, |r|/|b| {res / nb:.2e}, rel err vs FFT {err:.2e}")
        assert res <= tol * nb * 1.0001
        assert err < 1e-5
    assert okj and stu == 0
    assert itu < itj


# ------------------------------------------------------------- sial216

def test_sial216_fp64_mixed_learned(unet_weights):
    from sparc_amd import forces as F
    from sparc_amd import gpu
    from sparc_amd.sparc_io import occupations
    from sparc_amd.systems import scf_system
    st, p = scf_system("sial216")
    assert st.n_atom == 216 and p["nstates"] == 480 and p["smear_type"] == 0
    g = st.grid
    runs = {}
    with gpu.Context(g) as ctx:
        ctx.set_nonlocal(st.nloc)
        net = ctx.unet(seed=1)
        net.set_weights(unet_weights)
        # the *_tight runs solve every Poisson problem to 1e-10 instead of
        # TOL_POISSON = TOL_SCF / 100
        plan = [("f64", gpu.F64, None, None), ("mixed", gpu.F32, None, None),
                ("learned", gpu.F32, net, None), ("mixed_tight", gpu.F32, None, 1e-10),
                ("learned_tight", gpu.F32, net, 1e-10)]
        for tag, dt, pn, tolp in plan:
            _step(f"sial216 SCF {tag}")
            pp = (dict(p, dtype=dt) if tolp is None
                  else dict(p, dtype=dt, tol_poisson=tolp, maxit_poisson=10000))
            r = ctx.scf_run(st.b, st.rho_at, pp, rho_core=st.rho_core, poisson_net=pn,
                            keep_psi=True)
            occ = occupations(r["lambda_"], r["Efermi"], p["beta"], p["smear_type"])
            ints = ctx.nonlocal_integrals(r["psi"])
            fnl = F.nonlocal_forces_from_integrals(st, ints, occ, p["occfac"])
            floc, fxc = F.local_forces(ctx, st, r["phi"], r["vxc"])
            r["forces"] = F.symmetrize(floc + fnl + fxc)
            r.pop("psi").free()
            runs[tag] = r
        net.close()
    f = runs["f64"]
    assert f["converged"] == 1
    for tag in ("mixed", "learned"):
        r = runs[tag]
        de = abs(r["Etot"] - f["Etot"]) / st.n_atom
        dF = np.abs(r["forces"] - f["forces"]).max()
        print(f"sial216 {tag}: {r['scf_iters']} iters, {r['mean_iter_ms']:.1f} ms/iter (FP64 "
              f"{f['mean_iter_ms']:.1f}), dE/atom {de:.2e}, max dF {dF:.2e}, Poisson AAR "
              f"{r['poisson_iters']} (FP64 {f['poisson_iters']}), fallbacks {r['poisson_fallbacks']}")
        assert r["converged"] == 1
        assert de < 1e-5
        # the mixed run follows the FP64 trajectory; the learned run's Poisson
        # solves stop at other iterates within TOL_POISSON (2.2e-6 relative
        # residual here), which moves the forces of an SCF converged to
        # TOL_SCF = 2.2e-4 by ~1e-4 Ha/Bohr (1.0e-4 measured on 78^3,
        # 5.2e-5 on 80^3)
        assert dF < (1e-6 if tag == "mixed" else 3e-4)
    # with the Poisson solves converged far below that, the learned
    # preconditioner gives the forces of the Jacobi solves to 1e-6 Ha/Bohr
    mt, lt = runs["mixed_tight"], runs["learned_tight"]
    dFt = np.abs(lt["forces"] - mt["forces"]).max()
    det = abs(lt["Etot"] - mt["Etot"]) / st.n_atom
    print(f"sial216 Poisson tol 1e-10: learned vs Jacobi max dF {dFt:.2e}, dE/atom {det:.2e}, "
          f"Poisson AAR {lt['poisson_iters']} vs {mt['poisson_iters']}")
    assert mt["converged"] == 1 and lt["converged"] == 1
    assert dFt < 1e-6
    assert runs["learned"]["poisson_iters"] < f["poisson_iters"]
