"""Device SCF (sgpu_scf_run) against the FP64 oracle SCF on the 8-atom
silicon-like model (26^3 grid, 24 states, degree 20).

Bars: FP64 device run -- same iteration count, eigenvalues within 1e-6 Ha and
total energy within 1e-5 Ha/atom of the oracle (plus a much tighter check on
the first iterations, where both follow the same trajectory); mixed precision
(FP32 filter and projections) -- total energy within 1e-5 Ha/atom and
eigenvalues within 1e-5 Ha of the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    from oracle import oracle as orc
    from oracle.scf import OracleSCF
    from sparc_amd.model import default_params, silicon_model
    from sparc_amd.workloads import uniform_block
    orc.set_threads(16)
    sysm = silicon_model(1, 26)
    p = default_params(sysm)
    nd = sysm.grid.nd
    psi0 = uniform_block(nd, p["nstates"], p["psi_seed"], -0.5, 0.5)
    x0 = uniform_block(nd, 1, p["lanczos_seed"], 0.0, 1.0)
    ref = OracleSCF(sysm.grid, sysm.nloc, sysm.b, sysm.rho0, p, psi0, x0).run()
    return sysm, p, ref


def _run(sysm, p):
    from sparc_amd import gpu
    with gpu.Context(sysm.grid) as ctx:
        ctx.set_nonlocal(sysm.nloc)
        return ctx.scf_run(sysm.b, sysm.rho0, p)


def test_scf_fp64_matches_oracle(case):
    sysm, p, ref = case
    r = _run(sysm, p)
    print(f"\nFP64 SCF: iters {r['scf_iters']} (oracle {ref['scf_iters']}), Etot {r['Etot']:.10f} "
          f"(oracle {ref['Etot']:.10f}), mean iter {r['mean_iter_ms']:.1f} ms")
    assert r["converged"] == 1
    assert r["scf_iters"] == ref["scf_iters"]
    # first iterations: the eigenvectors after the first passes are far from
    # converged, so roundoff in the Lanczos bounds shows at the 1e-8 level
    k = min(6, r["scf_iters"])
    assert np.max(np.abs(r["iter_etot"][:k] - ref["iter_etot"][:k])) < 1e-6
    assert abs(r["Etot"] - ref["Etot"]) / sysm.n_atom < 1e-5
    assert np.max(np.abs(r["lambda_"] - ref["lambda_"])) < 1e-6
    assert abs(r["Efermi"] - ref["Efermi"]) < 1e-6
    rho_err = np.linalg.norm(r["rho"] - ref["rho"]) / np.linalg.norm(ref["rho"])
    assert rho_err < 1e-6
    for key in ("Eband", "Exc", "E1", "E2", "E3", "Entropy"):
        assert abs(r[key] - ref[key]) < 1e-5, key


def test_scf_mixed_precision_matches_oracle(case):
    from sparc_amd import gpu
    sysm, p, ref = case
    r = _run(sysm, dict(p, dtype=gpu.F32))
    de = abs(r["Etot"] - ref["Etot"]) / sysm.n_atom
    dl = np.max(np.abs(r["lambda_"] - ref["lambda_"]))
    print(f"\nmixed SCF: iters {r['scf_iters']} (oracle {ref['scf_iters']}), dE/atom {de:.2e}, "
          f"max dlambda {dl:.2e}")
    assert r["converged"] == 1
    assert de < 1e-5
    assert dl < 1e-5


def test_scf_rejects_bad_params(case):
    from sparc_amd import gpu
    sysm, p, _ = case
    with gpu.Context(sysm.grid) as ctx:
        ctx.set_nonlocal(sysm.nloc)
        for bad in (dict(nstates=10), dict(mix_history=9), dict(dtype=7), dict(beta=-1.0)):
            with pytest.raises(gpu.SparcGPUError):
                ctx.scf_run(sysm.b, sysm.rho0, dict(p, **bad))


@pytest.fixture(scope="module")
def sic():
    import json
    from pathlib import Path
    from oracle import oracle as orc
    from oracle.scf import OracleSCF
    from sparc_amd.systems import sic8
    from sparc_amd.workloads import uniform_block
    gold = json.loads((Path(__file__).parent / "golden" / "sic8_refout.json").read_text())
    orc.set_threads(16)
    st, p = sic8()
    nd = st.grid.nd
    psi0 = uniform_block(nd, p["nstates"], p["psi_seed"], -0.5, 0.5)
    x0 = uniform_block(nd, 1, p["lanczos_seed"], 0.0, 1.0)
    ref = OracleSCF(st.grid, st.nloc, st.b, st.rho_at, p, psi0, x0, rho_core=st.rho_core).run()
    return st, p, gold, ref


def _run_psp(st, p):
    from sparc_amd import gpu
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        return ctx.scf_run(st.b, st.rho_at, p, rho_core=st.rho_core)


def test_scf_sic_fp64_matches_reference_and_oracle(sic):
    """psp8 SiC (LDA_PW, NLCC): the FP64 device SCF reaches the reference's
    printed free energy and the oracle's eigenvalues."""
    st, p, gold, ref = sic
    r = _run_psp(st, p)
    e = r["Etot"] / st.n_atom
    print(f"\nSiC FP64: {r['scf_iters']} iters, E/atom {e:.10f} (reference "
          f"{gold['free_energy_per_atom']:.10f}, oracle {ref['Etot'] / st.n_atom:.10f})")
    assert r["converged"] == 1
    assert abs(e - gold["free_energy_per_atom"]) < 1e-6
    assert abs(r["Efermi"] - gold["fermi_level"]) < 1e-6
    assert abs(r["Exc"] - gold["exchange_correlation_energy"]) < 1e-5
    assert r["scf_iters"] == ref["scf_iters"]
    assert np.max(np.abs(r["lambda_"] - ref["lambda_"])) < 1e-6


def test_scf_sic_mixed_matches_reference(sic):
    from sparc_amd import gpu
    st, p, gold, ref = sic
    r = _run_psp(st, dict(p, dtype=gpu.F32))
    de = abs(r["Etot"] / st.n_atom - gold["free_energy_per_atom"])
    dl = np.max(np.abs(r["lambda_"] - ref["lambda_"]))
    print(f"\nSiC mixed: {r['scf_iters']} iters, dE/atom vs reference {de:.2e}, "
          f"max dlambda vs oracle {dl:.2e}")
    assert r["converged"] == 1
    assert de < 1e-5
    assert dl < 1e-5


@pytest.fixture(scope="module")
def si8u():
    """8-atom Si on 28^3 (a grid the U-Net's two halvings accept) and its
    FP64 Jacobi SCF, plus a U-Net trained on the device on (residual, error)
    pairs of Jacobi-AAR Poisson solves of the 64-atom cell (52^3: on 28^3 the
    quarter level is 7^3 and 400 steps do not train reliably), used across
    grids as in the benchmark.  A second network is trained for the Kerker
    solves (shift -kTF^2) on Jacobi-AAR solves whose right-hand sides are
    (Lap - 0.1 kTF^2) applied to the 64-atom SCF's first density changes."""
    from sparc_amd import gpu
    from sparc_amd.systems import scf_system
    gpu.load()
    s64, p64 = scf_system("si64")
    k2 = p64["kerker_ktf"] ** 2
    with gpu.Context(s64.grid) as ctx:
        ctx.set_nonlocal(s64.nloc)
        rhos = [s64.rho_at]
        for k in (1, 2, 3):
            rhos.append(ctx.scf_run(s64.b, s64.rho_at, dict(p64, max_scf=k, dtype=gpu.F32),
                                    rho_core=s64.rho_core)["rho"])
        net = ctx.unet(seed=1)
        for rho in (rhos[0], rhos[2]):
            net.harvest(4.0 * np.pi * (s64.b + rho), np.zeros(s64.grid.nd), tol=1e-8,
                        max_iter=3000, stride=2, max_pairs=120)
        hist = net.train(400, batch=4, lr=2e-3, seed=5)
        w = net.get_weights()
        net.close()
        knet = ctx.unet(seed=1)
        knet.set_shift(-k2)
        for a, b in zip(rhos[:-1], rhos[1:]):
            knet.harvest(ctx.Lap_vec_mult(b - a, -k2 * p64["kerker_thresh"]).ravel(),
                         np.zeros(s64.grid.nd), tol=1e-8, max_iter=3000, stride=2, max_pairs=60)
        khist = knet.train(400, batch=4, lr=2e-3, seed=5)
        wk = knet.get_weights()
        knet.close()
    st, p = scf_system("si8u")
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        rj = ctx.scf_run(st.b, st.rho_at, dict(p, dtype=gpu.F64), rho_core=st.rho_core)
    return st, p, rj, w, hist, wk, khist


def _learned_scf(st, p, w, wk=None):
    from sparc_amd import gpu
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        net = ctx.unet(seed=1)
        net.set_weights(w)
        knet = None
        if wk is not None:
            knet = ctx.unet(seed=1)
            knet.set_weights(wk)
            knet.set_shift(-p["kerker_ktf"] ** 2)
        return ctx.scf_run(st.b, st.rho_at, p, rho_core=st.rho_core, poisson_net=net,
                           kerker_net=knet)


def test_scf_learned_poisson_preconditioner(si8u):
    """The SCF with the trained U-Net as precond_fun of its Poisson solves
    converges to the Jacobi SCF's energy (the solves stop at the same relative
    residual; tolerance 1e-6 Ha/atom stated here) with fewer Poisson AAR
    iterations in total and no solve handed back to Jacobi."""
    from sparc_amd import gpu
    st, p, rj, w, hist, _, _ = si8u
    assert hist[-20:].mean() < 0.2 * hist[:5].mean()
    ru = _learned_scf(st, dict(p, dtype=gpu.F64), w)
    assert rj["converged"] and ru["converged"]
    assert ru["poisson_fallbacks"] == 0
    assert abs(ru["Etot"] - rj["Etot"]) / st.n_atom < 1e-6
    assert np.max(np.abs(ru["lambda_"] - rj["lambda_"])) < 1e-5
    assert ru["poisson_iters"] < rj["poisson_iters"], (ru["poisson_iters"], rj["poisson_iters"])


def test_scf_learned_preconditioner_falls_back_to_jacobi(si8u):
    """A preconditioner that makes AAR diverge (random weights of a large
    scale) does not derail the SCF: every solve that does not converge within
    maxit_poisson is redone with Jacobi-AAR from the same start, and the SCF
    reaches the Jacobi SCF's energy."""
    from sparc_amd import gpu
    st, p, rj, w, _, _, _ = si8u
    bad = np.random.default_rng(3).standard_normal(w.size).astype(np.float32) * 0.5
    ru = _learned_scf(st, dict(p, dtype=gpu.F64, maxit_poisson=300), bad)
    assert ru["converged"]
    assert ru["poisson_fallbacks"] > 0
    assert abs(ru["Etot"] - rj["Etot"]) / st.n_atom < 1e-6


def test_scf_learned_kerker_preconditioner(si8u):
    """Kerker_precond (src/mixing.c:462-527) with the Kerker-trained U-Net as
    precond_fun of its AAR solves: the SCF reaches the Jacobi SCF's energy
    (1e-6 Ha/atom) with fewer Kerker AAR iterations in total and no solve
    handed back to Jacobi."""
    from sparc_amd import gpu
    st, p, rj, w, _, wk, khist = si8u
    assert khist[-20:].mean() < 0.2 * khist[:5].mean()
    ru = _learned_scf(st, dict(p, dtype=gpu.F64), w, wk)
    assert ru["converged"]
    assert ru["kerker_fallbacks"] == 0
    assert abs(ru["Etot"] - rj["Etot"]) / st.n_atom < 1e-6
    assert ru["kerker_iters"] < rj["kerker_iters"], (ru["kerker_iters"], rj["kerker_iters"])


def test_scf_learned_kerker_falls_back_to_jacobi(si8u):
    """A Kerker preconditioner whose output is not finite (weights of 1e30
    overflow FP32 within the network) stops each solve at its first residual
    check; the solve is redone with Jacobi-AAR from the same start and the
    SCF still reaches the Jacobi SCF's energy."""
    from sparc_amd import gpu
    st, p, rj, w, _, wk, _ = si8u
    bad = np.full(wk.size, 1e30, np.float32)
    ru = _learned_scf(st, dict(p, dtype=gpu.F64), w, bad)
    assert ru["converged"]
    assert ru["kerker_fallbacks"] > 0
    assert abs(ru["Etot"] - rj["Etot"]) / st.n_atom < 1e-6


def test_unet_shift_must_not_be_positive(si8u):
    from sparc_amd import gpu
    st = si8u[0]
    with gpu.Context(st.grid) as ctx:
        net = ctx.unet(seed=1)
        with pytest.raises(gpu.SparcGPUError, match="shift"):
            net.set_shift(0.5)
        net.set_shift(-1.0)
        net.close()
