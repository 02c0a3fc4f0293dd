"""Reading the reference's .inpt / .ion inputs and writing .out / .eigen
outputs (sparc_amd/sparc_io.py): the reference's own SiC test input gives
the parameters the reference echoes for it and the same set-up as the
hand-built SiC system; unsupported physics is refused; the grid-from-ECUT
rule is checked against the FD symbol; outputs read back."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
SIC = GOLDEN / "sic8" / "SiC_orthogonal_quick_relax"


@pytest.fixture(scope="module")
def sic_case():
    from sparc_amd.sparc_io import load_case
    return load_case(SIC)


@pytest.fixture(scope="module")
def ref():
    return json.loads((GOLDEN / "sic8_refout.json").read_text())


NUMERIC = ["SMEARING", "TOL_SCF", "TOL_POISSON", "TOL_LANCZOS", "TOL_PSEUDOCHARGE", "TOL_PRECOND",
           "PRECOND_KERKER_KTF", "PRECOND_KERKER_THRESH", "MIXING_PARAMETER", "REFERENCE_CUTOFF"]
EXACT = ["FD_GRID", "FD_ORDER", "BC", "KPOINT_GRID", "SPIN_TYP", "ELEC_TEMP_TYPE",
         "EXCHANGE_CORRELATION", "NSTATES", "CHEB_DEGREE", "MAXIT_SCF", "MINIT_SCF", "MAXIT_POISSON",
         "POISSON_SOLVER", "MIXING_VARIABLE", "MIXING_PRECOND", "MIXING_HISTORY", "PULAY_FREQUENCY",
         "RHO_TRIGGER"]


def test_reference_input_gives_the_parameters_the_reference_echoes(sic_case, ref):
    """Grid from MESH_SPACING, NSTATES / CHEB_DEGREE / smearing / tolerance
    defaults: every effective value equals the one in the reference's own
    output for this input (printed with 3 significant digits for the
    tolerances)."""
    echo = ref["echo"]
    for k in EXACT:
        assert sic_case.echo[k] == echo[k], k
    for k in NUMERIC:
        a, b = float(sic_case.echo[k]), float(echo[k])
        assert abs(a - b) <= 5e-3 * abs(b), (k, a, b)
    assert sic_case.params["tol_precond"] == pytest.approx(0.34488 ** 2 * 1e-3, rel=1e-4)
    assert set(sic_case.ignored) >= {"RELAX_FLAG", "CALC_STRESS"}
    assert sic_case.print_forces and sic_case.ref_cutoff == 0.5


def test_reference_input_builds_the_validated_sic_system(sic_case):
    """The set-up read from the files equals systems.sic8() (the hand-built
    cell whose SCF matches the reference's energies): grid, atoms,
    pseudocharges, Esc, guess density, projectors and SCF parameters."""
    from sparc_amd.systems import sic8
    st, p = sic8()
    c = sic_case
    g, h = c.grid, st.grid
    assert (g.nx, g.ny, g.nz, g.order) == (h.nx, h.ny, h.nz, h.order)
    assert g.hx == pytest.approx(h.hx, rel=1e-14)
    assert c.setup.n_atom == st.n_atom == 8
    assert np.allclose(c.setup.b, st.b, rtol=0, atol=1e-12)
    assert np.allclose(c.setup.rho_at, st.rho_at, rtol=0, atol=1e-12)
    assert np.allclose(c.setup.rho_core, st.rho_core, rtol=0, atol=1e-12)
    assert c.setup.Esc == pytest.approx(st.Esc, abs=1e-10)
    assert np.array_equal(c.setup.nloc.grid_pos, st.nloc.grid_pos)
    assert np.allclose(c.setup.nloc.chi, st.nloc.chi, rtol=0, atol=1e-13)
    for k, v in p.items():
        if k in ("psi_seed", "lanczos_seed"):
            continue
        # beta: the reader applies the reference's default CONST_EH / 0.2 eV;
        # sic8() takes the smearing the reference printed (10 digits)
        rel = 1e-10 if k == "beta" else 1e-12
        assert c.params[k] == pytest.approx(v, rel=rel, abs=1e-15), k


def _write_case(tmp_path, inpt_extra="", ion_extra="", latvec="1 0 0\n0 1 0\n0 0 1"):
    psp = "14_Si_4_1.9_1.9_pbe_n_v1.0.psp8"
    (tmp_path / "c.inpt").write_text(
        f"LATVEC_SCALE: 10.26 10.26 10.26\nLATVEC:\n{latvec}\nFD_GRID: 16 16 16\n{inpt_extra}\n")
    (tmp_path / "c.ion").write_text(
        f"ATOM_TYPE: Si\nN_TYPE_ATOM: 2\nPSEUDO_POT: {psp}\nCOORD_FRAC:\n0 0 0\n0.25 0.25 0.25\n"
        f"{ion_extra}\n")
    return tmp_path / "c"


@pytest.mark.parametrize("extra,what", [
    ("EXCHANGE_CORRELATION: GGA_PBE", "EXCHANGE_CORRELATION"),
    ("KPOINT_GRID: 2 2 2", "KPOINT_GRID"),
    ("SPIN_TYP: 1", "SPIN_TYP"),
    ("BC: D D D", "BC"),
    ("MIXING_VARIABLE: potential", "MIXING_VARIABLE"),
    ("MD_FLAG: 1", "MD_FLAG"),
    ("EXX_ACC: 1", "not supported"),
])
def test_unsupported_physics_is_refused(tmp_path, extra, what):
    from sparc_amd.sparc_io import UnsupportedInput, load_case
    with pytest.raises(UnsupportedInput, match=what):
        load_case(_write_case(tmp_path, extra))


@pytest.mark.parametrize("extra,keys", [
    ("MESH_SPACING: 0.6", "FD_GRID, MESH_SPACING"),
    ("ECUT: 20", "FD_GRID, ECUT"),
    ("TOL_SCF: 1e-6\nACCURACY: 4", "TOL_SCF, ACCURACY"),
    ("SCF_ENERGY_ACC: 1e-5\nSCF_FORCE_ACC: 1e-3", "SCF_ENERGY_ACC, SCF_FORCE_ACC"),
    ("SMEARING: 0.01\nELEC_TEMP: 300", "SMEARING, ELEC_TEMP"),
])
def test_conflicting_keys_are_refused(tmp_path, extra, keys):
    """Keys the reference refuses together (src/readfiles.c:1101-1120): more
    than one way of giving the mesh, the SCF accuracy or the smearing."""
    from sparc_amd.sparc_io import load_case
    with pytest.raises(ValueError, match="cannot be simultaneously specified") as e:
        load_case(_write_case(tmp_path, extra))
    assert keys in str(e.value)


def test_non_orthogonal_cell_is_refused(tmp_path):
    from sparc_amd.sparc_io import UnsupportedInput, load_case
    with pytest.raises(UnsupportedInput, match="LATVEC"):
        load_case(_write_case(tmp_path, latvec="1 0 0\n0.5 0.866 0\n0 0 1"))


def test_defaults_of_a_minimal_input(tmp_path):
    """Two Si atoms, nothing but the cell and grid given: LDA_PZ, Gaussian
    0.2 eV smearing, TOL_SCF for 1e-5 Ha/atom, NSTATES = int(Ne/2 * 1.2 + 5),
    degree from the mesh, Cartesian and fractional coordinates alike."""
    from sparc_amd.sparc_io import load_case
    from sparc_amd.systems import mesh2chebdegree, tol_scf_for
    c = load_case(_write_case(tmp_path))
    p = c.params
    assert p["xc"] == 0 and p["smear_type"] == 1
    assert 1.0 / p["beta"] == pytest.approx(0.2 / 27.211386245988, rel=1e-12)
    assert p["tol_scf"] == pytest.approx(tol_scf_for(1e-5), rel=1e-12)
    assert p["tol_poisson"] == pytest.approx(p["tol_scf"] * 0.01)
    assert p["nstates"] == int(8 / 2 * 1.2 + 5)
    assert p["cheb_degree"] == mesh2chebdegree(10.26 / 16)
    # the same atoms in Cartesian coordinates
    ion = (tmp_path / "c.ion").read_text().replace("COORD_FRAC:\n0 0 0\n0.25 0.25 0.25",
                                                   "COORD:\n0 0 0\n2.565 2.565 2.565")
    (tmp_path / "c.ion").write_text(ion)
    c2 = load_case(tmp_path / "c")
    assert np.allclose(c2.setup.b, c.setup.b, atol=1e-13)


def test_ecut_grid_rule_against_the_fd_symbol():
    """ecut2h: h sqrt(2 Ecut) is the first k of the 1000-point grid of
    (0, pi] where the order-2n FD symbol (from the oracle's weights) departs
    from k^2 by more than 0.1."""
    from oracle import oracle as orc
    from sparc_amd.sparc_io import ecut2h
    for n in (2, 4, 6, 8):
        w = orc.fd_weights(n)
        for ecut in (10.0, 30.0):
            h = ecut2h(ecut, n)
            k = h * math.sqrt(2 * ecut)
            sym = lambda kk: -(w[0] + 2 * sum(w[p] * math.cos(kk * p) for p in range(1, n + 1)))
            assert abs(sym(k) - k * k) > 0.1
            km = k - math.pi / 1000
            assert km <= 1e-12 or abs(sym(km) - km * km) <= 0.1


def test_out_and_eigen_files_read_back(tmp_path, sic_case, ref):
    """write_out / write_eigen with the reference's final SiC energies: the
    energy block and SCF table read back; occupations sum to N_e."""
    from sparc_amd.sparc_io import read_out_energies, write_eigen, write_out
    from scipy.optimize import brentq
    from sparc_amd.sparc_io import occupations
    n = sic_case.setup.n_atom
    hist = ref["scf_history"]
    lam = np.linspace(-0.4, 0.5, sic_case.params["nstates"])
    beta, st = sic_case.params["beta"], sic_case.params["smear_type"]
    ef = brentq(lambda mu: 2 * occupations(lam, mu, beta, st).sum() - sic_case.params["nelec"],
                -1.0, 1.0, xtol=1e-14)
    res = {"iter_etot": np.array([h[1] * n for h in hist]), "iter_err": np.array([h[2] for h in hist]),
           "iter_ms": np.full(len(hist), 12.0), "scf_iters": len(hist),
           "Etot": ref["total_free_energy"], "Eband": ref["band_structure_energy"],
           "Exc": ref["exchange_correlation_energy"], "Entropy": ref["entropy_term"],
           "Efermi": ef, "lambda_": lam}
    res["forces"] = np.asarray(ref["forces_step1"])
    write_out(tmp_path / "s.out", sic_case, res)
    write_eigen(tmp_path / "s.eigen", sic_case, res)
    from sparc_amd.sparc_io import write_static
    write_static(tmp_path / "s.static", sic_case, res)
    st_text = (tmp_path / "s.static").read_text()
    fl = st_text.split("Atomic forces (Ha/Bohr):")[1].strip().splitlines()
    assert np.allclose([[float(v) for v in ln.split()] for ln in fl], ref["forces_step1"], rtol=1e-9)
    out_text = (tmp_path / "s.out").read_text()
    import re
    rms = float(re.search(r"RMS force\s*:\s*([-+0-9.E]+)", out_text).group(1))
    mx = float(re.search(r"Maximum force\s*:\s*([-+0-9.E]+)", out_text).group(1))
    assert rms == pytest.approx(ref["rms_force"], rel=1e-8)
    assert mx == pytest.approx(ref["max_force"], rel=1e-8)
    back = read_out_energies(tmp_path / "s.out")
    assert back["Free energy per atom"] == pytest.approx(ref["free_energy_per_atom"], abs=1e-10)
    assert back["Total free energy"] == pytest.approx(ref["total_free_energy"], rel=1e-10)
    assert back["Self and correction energy"] == pytest.approx(sic_case.setup.Esc, rel=1e-9)
    assert [(a, b) for a, b, _ in back["scf"]] == [(h[0], pytest.approx(h[1], abs=1e-10)) for h in hist]
    rows = (tmp_path / "s.eigen").read_text().splitlines()[5:]
    occ = np.array([float(r.split()[2]) for r in rows])
    assert len(rows) == len(lam)
    assert occ.sum() == pytest.approx(sic_case.params["nelec"], abs=1e-9)
