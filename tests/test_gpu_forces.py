"""Forces through the device: sgpu_nonlocal_integrals (the first-derivative
stencil of every orbital and the projections, images summed) against the
oracle's restatement in both projector layouts and precisions, and the
forces after the device SCF of the reference's SiC test input against the
forces the reference printed (SiC_orthogonal_quick_relax.refgeopt)."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import rel_err, small_case, split_images

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "sic8_refout.json"


@pytest.fixture(scope="module")
def gpu():
    from sparc_amd import gpu
    gpu.load()
    return gpu


@pytest.mark.parametrize("mode", ["brick", "sphere"])
@pytest.mark.parametrize("dtype", [0, 1])
def test_nonlocal_integrals_match_oracle(gpu, monkeypatch, mode, dtype):
    """Overlapping spheres, one atom split into two image entries, a grid
    with every dimension different: dV Chi' psi and Chi' d_d psi against the
    oracle (FP64 1e-12; FP32 input and arithmetic 2e-5)."""
    from oracle.forces import integrals
    monkeypatch.setenv("SPARC_AMD_NL_MODE", mode)
    g, veff, nl, rng = small_case(n=(16, 15, 14), n_sites=5, rc=2.6, seed=31, lmax=2)
    nl = split_images(nl, 1)
    X = rng.uniform(-0.5, 0.5, (g.nd, 21))
    with gpu.Context(g) as ctx:
        ctx.set_nonlocal(nl)
        b = ctx.block(21, dtype).upload(X)
        got = ctx.nonlocal_integrals(b)
    Xr = X if dtype == 0 else X.astype(np.float32).astype(np.float64)
    ref = integrals(g, nl, Xr)
    for k in range(4):
        assert rel_err(got[k], ref[k]) < (1e-12 if dtype == 0 else 2e-5), k


def test_nonlocal_integrals_psp_sic(gpu):
    """Real psp8 projectors (18 / 8 per atom, images across the faces)."""
    from oracle.forces import integrals
    from sparc_amd.systems import sic8
    st, _ = sic8()
    rng = np.random.default_rng(5)
    X = rng.uniform(-0.5, 0.5, (st.grid.nd, 12))
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        got = ctx.nonlocal_integrals(ctx.block(12, gpu.F64).upload(X))
    ref = integrals(st.grid, st.nloc, X)
    for k in range(4):
        assert rel_err(got[k], ref[k]) < 1e-12, k


@pytest.mark.parametrize("name", ["sic8", "si8"])
def test_local_forces_match_oracle(gpu, name):
    """sgpu_local_forces (local + NLCC terms) against the oracle's box loop on
    the same fields (smooth random phi / Vxc), SiC (two types, boxes reaching
    across every face) and Si: 1e-11 Ha/Bohr."""
    from oracle import forces as O
    from sparc_amd import forces as F
    from sparc_amd.systems import sic8, scf_system
    st = sic8()[0] if name == "sic8" else scf_system("si8")[0]
    g = st.grid
    rng = np.random.default_rng(8)
    k = [2 * np.pi * np.arange(n) / n for n in (g.nx, g.ny, g.nz)]
    Z, Y, X = np.meshgrid(k[2], k[1], k[0], indexing="ij")
    phi = (np.sin(X + 0.3) * np.cos(2 * Y) + 0.5 * np.sin(Z - Y)).ravel() + 0.01 * rng.standard_normal(g.nd)
    vxc = (-0.3 + 0.1 * np.cos(X) * np.sin(Y + Z)).ravel()
    with gpu.Context(g) as ctx:
        floc, fxc = F.local_forces(ctx, st, phi, vxc)
    ref_loc, ref_xc = O.local_forces(st, phi), O.xc_forces(st, vxc)
    assert np.abs(floc - ref_loc).max() < 1e-11, np.abs(floc - ref_loc).max()
    assert np.abs(fxc - ref_xc).max() < 1e-11, np.abs(fxc - ref_xc).max()
    assert np.abs(ref_loc).max() > 1e-3 and np.abs(ref_xc).max() > 1e-5


@pytest.mark.parametrize("dtype,tol", [(0, 5e-6), (1, 2e-5)])
def test_scf_forces_match_reference_sic(gpu, dtype, tol):
    """Device SCF (FP64, and mixed precision) of the SiC input, then the
    forces: within 5e-6 Ha/Bohr (FP64) / 2e-5 (mixed) of the forces the
    reference printed at this geometry; the three parts sum to the total
    after the mean is removed."""
    from sparc_amd import forces as F
    from sparc_amd.systems import sic8
    gold = json.loads(GOLDEN.read_text())
    st, p = sic8()
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        r = ctx.scf_run(st.b, st.rho_at, dict(p, dtype=dtype), rho_core=st.rho_core, keep_psi=True)
        assert r["converged"]
        f = F.scf_forces(ctx, st, p, r)
    ref = np.asarray(gold["forces_step1"])
    assert np.abs(f["total"] - ref).max() < tol, np.abs(f["total"] - ref).max()
    assert np.allclose(F.symmetrize(f["local"] + f["nonlocal"] + f["xc"]), f["total"], atol=1e-15)


def test_scf_rejects_psi_out_of_wrong_shape(gpu):
    from sparc_amd.systems import sic8
    st, p = sic8()
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        bad = ctx.block(p["nstates"] + 1, gpu.F64)
        from sparc_amd.gpu import SCFParams
        import ctypes as C
        with pytest.raises(gpu.SparcGPUError, match="psi_out"):
            q = dict(p, dtype=gpu.F64)
            pp = SCFParams(**q)
            pp.psi_out = bad.h.value
            lam = np.zeros(p["nstates"])
            res = gpu.SCFResult()
            ctx._check(ctx.lib.sgpu_scf_run(ctx.h, C.byref(pp), gpu._ptr(np.ascontiguousarray(st.b)),
                                            gpu._ptr(np.ascontiguousarray(st.rho_at)), None,
                                            gpu._ptr(lam), None, None, None, None, None, C.byref(res)))
