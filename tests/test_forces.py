"""Atomic forces (SPARC Calculate_EGS_Forces, src/forces.c): the oracle SCF
of the reference's own SiC test input followed by the oracle's force terms
(oracle/forces.py: local, NLCC and nonlocal)
reproduces the forces the reference printed at that geometry
(SiC_orthogonal_quick_relax.refgeopt, RELAXSTEP 1); the first-derivative
weights are checked against the derivative of a plane wave."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "sic8_refout.json"


@pytest.mark.parametrize("order", [4, 8, 12])
def test_first_derivative_weights_against_plane_wave(order):
    """sum_p w_p (e^{ikph} - e^{-ikph}) = i k + O((kh)^(order+1)): at kh =
    0.3 the order-2n stencil's error falls below (kh)^(2n+1) / h, and both
    the host weights and the oracle's agree."""
    from oracle.forces import d1_weights, fd_d1
    h = 0.4
    w = fd_d1(order, h)
    assert np.allclose(w, d1_weights(order, h), rtol=1e-15, atol=0)
    for kh in (0.05, 0.3):
        k = kh / h
        sym = sum(2.0 * w[p] * math.sin(p * kh) for p in range(1, len(w)))
        assert abs(sym - k) < (kh ** (order + 1)) / h + 1e-13
    # a plane wave on a periodic grid: the stencil maps sin(kx) to
    # (sum_p 2 w_p sin(p k h)) cos(kx) exactly, along each axis
    from oracle.forces import grad_periodic
    n = 32
    x = np.arange(n) * h
    k = 2 * math.pi * 3 / (n * h)
    sym = sum(2.0 * w[p] * math.sin(p * k * h) for p in range(1, len(w)))
    for axis in range(3):
        shape = [1, 1, 1]
        shape[axis] = n
        f = np.broadcast_to(np.sin(k * x).reshape(shape), (n, n, n)).copy()
        g = grad_periodic(f, w, axis)
        exact = np.broadcast_to(sym * np.cos(k * x).reshape(shape), (n, n, n))
        assert np.abs(g - exact).max() < 1e-12 * k


def test_oracle_scf_forces_match_reference_sic():
    """SiC (4 Si + 4 C, LDA-PZ, psp8 with NLCC, 26^3): the converged oracle
    SCF's forces equal the reference's printed forces within 5e-6 Ha/Bohr
    (its SCF stopped at error 8.6e-7; this SCF reproduces the same energies
    to 1e-7 Ha/atom), and the printed "RMS" (mean |F_J|) and maximum force within 1e-6."""
    from oracle import oracle as orc
    from oracle import forces as O
    from oracle.forces import nonlocal_forces
    from oracle.scf import OracleSCF
    from sparc_amd import forces as F
    from sparc_amd.sparc_io import occupations
    from sparc_amd.systems import sic8
    from sparc_amd.workloads import uniform_block
    gold = json.loads(GOLDEN.read_text())
    orc.set_threads(8)
    st, p = sic8()
    pos = np.concatenate([np.asarray(t.positions) for t in st.types])
    assert np.allclose(pos, gold["positions_step1"], atol=1e-12)
    nd = st.grid.nd
    psi0 = uniform_block(nd, p["nstates"], p["psi_seed"], -0.5, 0.5)
    x0 = uniform_block(nd, 1, p["lanczos_seed"], 0.0, 1.0)
    s = OracleSCF(st.grid, st.nloc, st.b, st.rho_at, p, psi0, x0, rho_core=st.rho_core)
    r = s.run()
    assert r["converged"]
    occ = occupations(r["lambda_"], r["Efermi"], p["beta"], p["smear_type"])
    floc = O.local_forces(st, s.phi)
    fxc = O.xc_forces(st, s.vxc)
    fnl = nonlocal_forces(st.grid, st.nloc, F.atom_rows(st), st.n_atom, r["X"], occ, p["occfac"])
    f = F.symmetrize(floc + fnl + fxc)
    ref = np.asarray(gold["forces_step1"])
    assert np.abs(f - ref).max() < 5e-6, np.abs(f - ref).max()
    norms = np.linalg.norm(f, axis=1)
    # the value printed as "RMS force" is the mean of |F_J|
    # (src/electronicGroundState.c:232-240)
    assert abs(norms.mean() - gold["rms_force"]) < 1e-6
    assert abs(norms.max() - gold["max_force"]) < 1e-6
    # every term matters at this tolerance: dropping one moves some force by far more
    for part in (floc, fnl, fxc):
        assert np.abs(F.symmetrize(floc + fnl + fxc - part) - ref).max() > 1e-4


@pytest.mark.parametrize("n,bc", [((13, 11, 9), (0, 0, 0)), ((16, 15, 14), (0, 1, 0)),
                                  ((9, 12, 10), (1, 1, 1))])
def test_oracle_gradient_matches_reference_gradient(n, bc):
    """The oracle's first derivative (periodic axes wrapped; the device
    kernel follows the same rule) equals SPARC's own compiled
    Gradient_vec_dir (src/gradVecRoutines.c:61-316, oracle/_ref) bit for bit
    on periodic grids, and on Dirichlet axes where the reference treats
    values outside the domain as zero."""
    from oracle import oracle as O
    from oracle.forces import d1_weights
    from sparc_amd.workloads import Grid
    if O.ref_lib() is None:
        pytest.skip("reference stencil (oracle/_ref) not built here")
    g = Grid(*n, 0.4, 0.45, 0.5, *bc)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((g.nd, 3))
    for axis in range(3):
        ref = O.ref_gradient(g, x, axis)
        h = (g.hx, g.hy, g.hz)[axis]
        w = d1_weights(g.order, h)
        V = x.T.reshape(3, g.nz, g.ny, g.nx)
        ax = 3 - axis
        nax = V.shape[ax]
        out = np.zeros_like(V)
        per = (g.bcx, g.bcy, g.bcz)[axis] == 0
        for p in range(1, len(w)):
            if per:
                a, b = np.roll(V, -p, axis=ax), np.roll(V, p, axis=ax)
            else:
                a = np.zeros_like(V); b = np.zeros_like(V)
                sl = [slice(None)] * 4
                sl[ax] = slice(0, nax - p); dst = tuple(sl)
                sl[ax] = slice(p, nax); src = tuple(sl)
                a[dst] = V[src]
                b[src] = V[dst]
            out += w[p] * (a - b)
        got = out.reshape(3, -1).T
        if per:
            from oracle.forces import gradient
            assert np.array_equal(gradient(g, x, axis), ref)
        assert np.array_equal(got, ref), axis
