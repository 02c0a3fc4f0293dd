"""SPARC's input and output files for the device SCF.

``read_inpt`` / ``read_ion`` read the reference's ``<name>.inpt`` and
``<name>.ion`` files (read_input src/readfiles.c:132-1460, read_ion
:1462-1700): ``KEY: values`` lines, ``#`` comments, keys case-insensitive;
atom types as ``ATOM_TYPE`` / ``N_TYPE_ATOM`` / ``PSEUDO_POT`` /
``COORD`` or ``COORD_FRAC`` blocks.  ``load_case`` turns them into the
grid, the atom set-up (pseudocharges, Esc, guess and core densities,
projectors) and the SCF parameters with the reference's defaults
(src/initialization.c:677-989 for the input defaults, :1716 NSTATES,
:1840-1856 smearing, :1930-1938 grid from MESH_SPACING / ECUT, :2236-2248
Chebyshev degree, :2269 simple-mixing parameter, :2509-2611 tolerances).
``write_out`` / ``write_eigen`` write the ``.out`` summary and the ``.eigen``
file in the reference's layout (src/initialization.c:3705-3830 input
echo, src/electronicGroundState.c:143-160,655-665,1015 SCF table and
energies, src/printing.c:225-430 eigenvalues).

Supported: the case the device path runs -- orthogonal cell (LATVEC a
scaled identity), periodic in x, y and z, Gamma point, spin-unpolarized,
LDA_PZ / LDA_PW, density mixing with the Kerker preconditioner.  Any other
physics raises ``UnsupportedInput`` naming the key; keys that only steer
output, relaxation, MD or forces are recorded in ``Case.ignored``.
"""
from __future__ import annotations

import math
import re
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .atoms import AtomType, setup_atoms
from .psp import PSP_DIR, read_psp8
from .systems import mesh2chebdegree
from .workloads import Grid

CONST_EH = 27.211386245988       # Hartree in eV (src/include/isddft.h:68)
CONST_KB = 3.1668115634556e-6    # Boltzmann constant in Ha/K (src/include/isddft.h:65)

# keys of the reference that steer output, restarts, relaxation, MD,
# forces, stress or parallelisation only: read and ignored here
_IGNORED = {
    "NP_SPIN_PARAL", "NP_KPOINT_PARAL", "NP_BAND_PARAL", "NP_DOMAIN_PARAL", "NP_DOMAIN_PHI_PARAL",
    "EIG_SERIAL_MAXNS", "EIG_PARAL_BLKSZ", "EIG_PARAL_ORFAC", "EIG_PARAL_MAXNP",
    "PRINT_ATOMS", "PRINT_VELS", "PRINT_EIGEN", "PRINT_DENSITY", "PRINT_RESTART", "PRINT_RESTART_FQ",
    "PRINT_RELAXOUT", "PRINT_MDOUT", "PRINT_ORBITAL", "PRINT_ENERGY_DENSITY", "VERBOSITY",
    "OUTPUT_FILE", "CALC_STRESS", "CALC_PRES", "RELAX_FLAG", "RELAX_METHOD", "RELAX_NITER",
    "TOL_RELAX", "L_HISTORY", "L_FINIT_STP", "L_MAXMOV", "L_AUTOSCALE", "L_LINEOPT", "L_ICURV",
    "NLCG_SIGMA", "FIRE_DT", "FIRE_MASS", "FIRE_MAXMOV", "RESTART_FLAG", "TWTIME", "FIX_RAND",
    "CHEFSI_BOUND_FLAG", "NUM_CHEFSI", "STANDARD_EIGEN", "POISSON_SOLVER", "KPOINT_SHIFT",
}


class UnsupportedInput(ValueError):
    """The input asks for physics the device SCF does not run."""


def _tokens(path: Path) -> list[list[str]]:
    out = []
    for line in Path(path).read_text().splitlines():
        line = line.split("#", 1)[0].strip()
        if line:
            out.append(line.split())
    return out


def read_inpt(path) -> dict[str, list[str]]:
    """``KEY: v1 v2 ...`` lines of a .inpt file -> {KEY (upper case): [values]};
    LATVEC takes the three following lines."""
    lines = _tokens(Path(path))
    keys: dict[str, list[str]] = {}
    i = 0
    while i < len(lines):
        head = lines[i][0]
        if not head.endswith(":"):
            raise ValueError(f"{path}: cannot read line {' '.join(lines[i])!r}")
        key = head[:-1].upper()
        vals = lines[i][1:]
        if key == "LATVEC":
            vals = [v for row in lines[i + 1:i + 4] for v in row]
            i += 3
        keys[key] = vals
        i += 1
    return keys


@dataclass
class IonType:
    name: str
    psp: str                   # PSEUDO_POT as written
    coords: np.ndarray         # [n, 3], fractional if frac else Bohr
    frac: bool


def read_ion(path) -> list[IonType]:
    """ATOM_TYPE blocks of a .ion file (read_ion, src/readfiles.c:1462)."""
    lines = _tokens(Path(path))
    types: list[dict] = []
    i = 0
    while i < len(lines):
        key = lines[i][0].rstrip(":").upper()
        vals = lines[i][1:]
        if key == "ATOM_TYPE":
            types.append({"name": vals[0], "n": None, "psp": None, "coords": None, "frac": False})
        elif not types:
            raise ValueError(f"{path}: {key} before the first ATOM_TYPE")
        elif key == "N_TYPE_ATOM":
            types[-1]["n"] = int(vals[0])
        elif key == "PSEUDO_POT":
            types[-1]["psp"] = vals[0]
        elif key in ("COORD", "COORD_FRAC"):
            n = types[-1]["n"]
            if n is None:
                raise ValueError(f"{path}: {key} of {types[-1]['name']} before N_TYPE_ATOM")
            rows = [[float(v) for v in lines[i + 1 + k][:3]] for k in range(n)]
            types[-1]["coords"] = np.asarray(rows, np.float64)
            types[-1]["frac"] = key == "COORD_FRAC"
            i += n
        elif key in ("RELAX", "SPIN"):
            n = types[-1]["n"] or 0
            if key == "SPIN":
                raise UnsupportedInput(f"{path}: SPIN (spin-polarized atoms) is not supported")
            i += n                         # relaxation flags: not used by the SCF
        elif key == "ATOMIC_MASS":
            pass
        else:
            raise ValueError(f"{path}: unknown key {key}")
        i += 1
    out = []
    for t in types:
        if t["psp"] is None or t["coords"] is None:
            raise ValueError(f"{path}: atom type {t['name']} lacks PSEUDO_POT or coordinates")
        out.append(IonType(t["name"], t["psp"], t["coords"], t["frac"]))
    return out


def ecut2h(ecut: float, fdn: int) -> float:
    """Mesh size equivalent to a plane-wave cutoff (Ecut2h,
    src/initialization.c:3094-3141): the first k on a 1000-point grid of
    (0, pi] where the FD symbol departs from k^2 by more than 0.1, over
    sqrt(2 Ecut)."""
    n = fdn
    w = np.zeros(n + 1)
    for p in range(1, n + 1):
        fr = math.factorial(n) ** 2 / (math.factorial(n - p) * math.factorial(n + p))
        w[p] = (2 * (p % 2) - 1) * 2 * fr / (p * p)
        w[0] -= 2.0 / (p * p)
    kk = (np.arange(1000) + 1) * (math.pi / 1000)
    y = -w[0] - 2 * sum(w[p] * np.cos(kk * p) for p in range(1, n + 1))
    bad = np.nonzero(np.abs(y - kk * kk) > 0.1)[0]
    k_cut = kk[bad[0]] if bad.size else kk[0]
    return float(k_cut / math.sqrt(2.0 * ecut))


@dataclass
class Case:
    name: str
    grid: Grid
    setup: object              # AtomSetup
    params: dict               # sgpu_scf_params fields
    echo: dict                 # effective input values in the .out echo order
    type_names: list
    type_counts: list
    psp_files: list
    ignored: list = field(default_factory=list)
    print_forces: bool = False  # PRINT_FORCES: forces after the SCF
    ref_cutoff: float = 0.5     # REFERENCE_CUTOFF of the pseudocharge correction


def _f(keys, k, default=None, idx=0):
    return float(keys[k][idx]) if k in keys else default


def _i(keys, k, default=None, idx=0):
    return int(keys[k][idx]) if k in keys else default


def _resolve_psp(ion_dir: Path, psp: str) -> Path:
    p = Path(psp)
    for cand in ((ion_dir / p) if not p.is_absolute() else p, PSP_DIR / p.name):
        if cand.exists():
            return cand
    raise FileNotFoundError(f"pseudopotential {psp} not found next to the .ion file or in {PSP_DIR}")


def load_case(base, psp_dir=None) -> Case:
    """``base`` = path without extension of a ``.inpt`` / ``.ion`` pair."""
    base = Path(base)
    keys = read_inpt(base.with_suffix(".inpt"))
    ions = read_ion(base.with_suffix(".ion"))
    ignored = sorted(k for k in keys if k in _IGNORED)

    def need(cond, what):
        if not cond:
            raise UnsupportedInput(f"{base.name}: {what}")

    # physics the device path runs
    kp = [int(v) for v in keys.get("KPOINT_GRID", ["1", "1", "1"])]
    need(kp == [1, 1, 1], f"KPOINT_GRID {kp}: only the Gamma point is supported")
    need(_i(keys, "SPIN_TYP", 0) == 0, "SPIN_TYP: only spin-unpolarized runs are supported")
    xc = keys.get("EXCHANGE_CORRELATION", ["LDA_PZ"])[0].upper()
    need(xc in ("LDA_PZ", "LDA_PW"), f"EXCHANGE_CORRELATION {xc}: only LDA_PZ / LDA_PW")
    bc = [v.upper() for v in keys.get("BC", [])]
    if "BOUNDARY_CONDITION" in keys:
        need(keys["BOUNDARY_CONDITION"][0] == "2", "BOUNDARY_CONDITION: only 2 (periodic)")
        bc = ["P", "P", "P"]
    need(bc in ([], ["P", "P", "P"]), f"BC {' '.join(bc)}: only P P P")
    need(keys.get("MIXING_VARIABLE", ["density"])[0].lower() == "density",
         "MIXING_VARIABLE: only density mixing")
    need(keys.get("MIXING_PRECOND", ["kerker"])[0].lower() == "kerker",
         "MIXING_PRECOND: only kerker")
    need(_i(keys, "PULAY_RESTART", 0) == 0, "PULAY_RESTART: not supported")
    for k in ("MD_FLAG", "OFDFT_FLAG", "SQ_FLAG", "MLFF_FLAG", "D3_FLAG", "HUBBARD_FLAG"):
        need(_i(keys, k, 0) == 0, f"{k}: not supported")
    unknown = [k for k in keys if k not in _IGNORED and k not in _KNOWN]
    need(not unknown, f"keys not supported: {', '.join(sorted(unknown))}")
    # keys the reference refuses in combination (src/readfiles.c:1101-1120)
    for group in (("FD_GRID", "MESH_SPACING", "ECUT"),
                  ("TOL_SCF", "ACCURACY", "SCF_ENERGY_ACC", "SCF_FORCE_ACC"),
                  ("SMEARING", "ELEC_TEMP")):
        given = [k for k in group if k in keys]
        if len(given) > 1:
            raise ValueError(f"{base.name}: {', '.join(given)} cannot be simultaneously specified")

    # cell: LATVEC must be a scaled identity (orthogonal axes along x, y, z)
    lat = np.asarray([float(v) for v in keys.get("LATVEC", ["1", "0", "0", "0", "1", "0", "0", "0", "1"])])
    lat = lat.reshape(3, 3)
    norms = np.linalg.norm(lat, axis=1)
    need(np.allclose(lat / norms[:, None], np.eye(3), atol=1e-12), "LATVEC: only orthogonal cells along x, y, z")
    if "LATVEC_SCALE" in keys:
        L = np.asarray([float(v) for v in keys["LATVEC_SCALE"]]) * norms
    elif "CELL" in keys:
        L = np.asarray([float(v) for v in keys["CELL"]])
    else:
        raise ValueError(f"{base.name}: LATVEC_SCALE or CELL is required")
    order = _i(keys, "FD_ORDER", 12)
    fdn = order // 2
    if "FD_GRID" in keys:
        n = [int(v) for v in keys["FD_GRID"]]
    elif "MESH_SPACING" in keys:
        h = float(keys["MESH_SPACING"][0])
        n = [max(math.ceil(l / h), fdn) for l in L]
    elif "ECUT" in keys:
        h = ecut2h(float(keys["ECUT"][0]), fdn)
        n = [max(math.ceil(l / h), fdn) for l in L]
    else:
        raise ValueError(f"{base.name}: FD_GRID, MESH_SPACING or ECUT is required")
    g = Grid(n[0], n[1], n[2], L[0] / n[0], L[1] / n[1], L[2] / n[2], order=order)
    if abs(g.hx - g.hy) < 1e-12 and abs(g.hy - g.hz) < 1e-12:
        h_eff = g.hx
    else:
        h_eff = math.sqrt(3.0 / (1 / g.hx ** 2 + 1 / g.hy ** 2 + 1 / g.hz ** 2))

    # tolerances (src/initialization.c:2509-2611)
    tol_scf = _f(keys, "TOL_SCF", -1.0)
    if tol_scf < 0:
        if _i(keys, "RELAX_FLAG", 0):
            fa = _f(keys, "TOL_RELAX", 5e-4) / 5
            tol_scf = math.exp((math.log(fa) - 1.468) / 1.025)
        else:
            acc = _i(keys, "ACCURACY", -1)
            fa = 10.0 ** -(acc + 1.0) if acc >= 0 else _f(keys, "SCF_FORCE_ACC", -1.0)
            ea = _f(keys, "SCF_ENERGY_ACC", -1.0) if fa < 0 else -1.0
            if fa < 0 and ea < 0:
                ea = 1e-5
            tol_scf = (math.exp((math.log(ea) - 1.165) / 1.502) if ea > 0
                       else math.exp((math.log(fa) - 1.368) / 1.025))
    tol_poisson = _f(keys, "TOL_POISSON", tol_scf * 0.01)
    tol_pseudocharge = _f(keys, "TOL_PSEUDOCHARGE", tol_scf * 0.001)
    tol_precond = _f(keys, "TOL_PRECOND", _f(keys, "TOL_KERKER", h_eff * h_eff * 1e-3))

    # atoms
    types, names, counts, files = [], [], [], []
    for t in ions:
        pos = t.coords * L[None, :] if t.frac else t.coords
        f = _resolve_psp(base.parent, t.psp) if psp_dir is None else Path(psp_dir) / Path(t.psp).name
        types.append(AtomType(read_psp8(f), np.mod(pos, L[None, :]), name=t.name))
        names.append(t.name)
        counts.append(len(t.coords))
        files.append(str(t.psp))
    st = setup_atoms(g, types, tol_pseudocharge=tol_pseudocharge,
                     ref_cutoff=_f(keys, "REFERENCE_CUTOFF", 0.5))
    # occupations use the pseudocharges' total (-int b, as the validated
    # systems do); the default band count uses the exact valence count
    net = _f(keys, "NET_CHARGE", 0.0)
    nelec = st.nelec - net
    zval = sum(t.psp.zion * len(t.positions) for t in types) - net
    nstates = _i(keys, "NSTATES", int((zval / 2) * 1.2 + 5))

    # smearing (src/readfiles.c:314-349, src/initialization.c:1840-1856)
    smear_type = 1
    if "ELEC_TEMP_TYPE" in keys:
        v = keys["ELEC_TEMP_TYPE"][0].lower()
        if v in ("fd", "fermi-dirac"):
            smear_type = 0
        elif v != "gaussian":
            raise ValueError(f"{base.name}: ELEC_TEMP_TYPE {v}")
    if "SMEARING" in keys:
        beta = 1.0 / float(keys["SMEARING"][0])
    elif "ELEC_TEMP" in keys:
        beta = 1.0 / (CONST_KB * float(keys["ELEC_TEMP"][0]))
    else:
        beta = CONST_EH / (0.2 if smear_type == 1 else 0.1)
    mix = _f(keys, "MIXING_PARAMETER", 0.3)
    params = dict(
        nstates=nstates, cheb_degree=_i(keys, "CHEB_DEGREE", mesh2chebdegree(h_eff)),
        rho_trigger=_i(keys, "RHO_TRIGGER", 4), max_scf=_i(keys, "MAXIT_SCF", 100),
        min_scf=_i(keys, "MINIT_SCF", 2), tol_scf=tol_scf, tol_poisson=tol_poisson,
        maxit_poisson=_i(keys, "MAXIT_POISSON", 3000), tol_lanczos=_f(keys, "TOL_LANCZOS", 1e-2),
        beta=beta, smear_type=smear_type, nelec=nelec, occfac=2.0,
        mix_history=_i(keys, "MIXING_HISTORY", 7), pulay_freq=_i(keys, "PULAY_FREQUENCY", 1),
        mix_beta=mix, mix_omega=_f(keys, "MIXING_PARAMETER_SIMPLE", mix),
        kerker_ktf=_f(keys, "PRECOND_KERKER_KTF", 1.0),
        kerker_thresh=_f(keys, "PRECOND_KERKER_THRESH", 0.1), tol_precond=tol_precond,
        Esc=st.Esc, xc=1 if xc == "LDA_PW" else 0, dtype=0, psi_seed=1234, lanczos_seed=4321)
    echo = {
        "LATVEC_SCALE": " ".join(f"{v:.15g}" for v in L),
        "FD_GRID": f"{g.nx} {g.ny} {g.nz}", "FD_ORDER": str(order), "BC": "P P P",
        "KPOINT_GRID": "1 1 1", "SPIN_TYP": "0",
        "ELEC_TEMP_TYPE": "Gaussian" if smear_type == 1 else "Fermi-Dirac",
        "SMEARING": f"{1.0 / beta:.10g}", "EXCHANGE_CORRELATION": xc, "NSTATES": str(nstates),
        "CHEB_DEGREE": str(params["cheb_degree"]), "MAXIT_SCF": str(params["max_scf"]),
        "MINIT_SCF": str(params["min_scf"]), "MAXIT_POISSON": str(params["maxit_poisson"]),
        "TOL_SCF": f"{tol_scf:.2E}", "POISSON_SOLVER": "AAR", "TOL_POISSON": f"{tol_poisson:.2E}",
        "TOL_LANCZOS": f"{params['tol_lanczos']:.2E}", "TOL_PSEUDOCHARGE": f"{tol_pseudocharge:.2E}",
        "MIXING_VARIABLE": "density", "MIXING_PRECOND": "kerker", "TOL_PRECOND": f"{tol_precond:.2E}",
        "PRECOND_KERKER_KTF": f"{params['kerker_ktf']:.10g}",
        "PRECOND_KERKER_THRESH": f"{params['kerker_thresh']:.10g}",
        "MIXING_PARAMETER": f"{mix:.10g}", "MIXING_HISTORY": str(params["mix_history"]),
        "PULAY_FREQUENCY": str(params["pulay_freq"]),
        "REFERENCE_CUTOFF": f"{_f(keys, 'REFERENCE_CUTOFF', 0.5):.10g}",
        "RHO_TRIGGER": str(params["rho_trigger"]),
    }
    st.meta.update(nelec=nelec)
    return Case(base.name, g, st, params, echo, names, counts, files, ignored,
                print_forces=_i(keys, "PRINT_FORCES", 0) == 1,
                ref_cutoff=_f(keys, "REFERENCE_CUTOFF", 0.5))


_KNOWN = {
    "LATVEC_SCALE", "CELL", "LATVEC", "BC", "BOUNDARY_CONDITION", "FD_ORDER", "FD_GRID",
    "MESH_SPACING", "ECUT", "KPOINT_GRID", "SPIN_TYP", "ELEC_TEMP_TYPE", "SMEARING", "ELEC_TEMP",
    "CHEB_DEGREE", "RHO_TRIGGER", "NSTATES", "NET_CHARGE", "MAXIT_SCF", "MINIT_SCF", "MAXIT_POISSON",
    "ACCURACY", "SCF_FORCE_ACC", "SCF_ENERGY_ACC", "TOL_SCF", "TOL_POISSON", "TOL_LANCZOS",
    "TOL_PSEUDOCHARGE", "TOL_KERKER", "TOL_PRECOND", "PRECOND_KERKER_KTF", "PRECOND_KERKER_THRESH",
    "REFERENCE_CUTOFF", "MIXING_VARIABLE", "MIXING_PRECOND", "MIXING_HISTORY", "MIXING_PARAMETER",
    "MIXING_PARAMETER_SIMPLE", "PULAY_FREQUENCY", "PULAY_RESTART", "EXCHANGE_CORRELATION",
    "MD_FLAG", "OFDFT_FLAG", "SQ_FLAG", "MLFF_FLAG", "D3_FLAG", "HUBBARD_FLAG", "PRINT_FORCES",
}


def occupations(lam: np.ndarray, efermi: float, beta: float, smear_type: int) -> np.ndarray:
    """Occupation per state (without the spin factor): Fermi-Dirac
    1/(1+exp(beta(l-mu))) or Gaussian erfc(beta(l-mu))/2 (smearing_function,
    src/occupation.c)."""
    from scipy.special import erfc, expit
    x = beta * (np.asarray(lam) - efermi)
    return expit(-x) if smear_type == 0 else 0.5 * erfc(x)


def write_out(path, case: Case, result: dict, title: str = "device SCF (MI355X)") -> None:
    """The .out summary in the reference's layout: input echo, cell, SCF
    table (free energy per atom, SCF error, wall time per iteration) and the
    energy block."""
    g = case.grid
    L = (g.hx * g.nx, g.hy * g.ny, g.hz * g.nz)
    n_atom = case.setup.n_atom
    bar = "*" * 75
    lines = [bar, f"*{title:^73s}*", f"*{'Start time: ' + time.ctime():^73s}*", bar,
             f"{'Input parameters':^75s}".rstrip(), bar]
    for k, v in case.echo.items():
        if k == "LATVEC_SCALE":
            lines += [f"LATVEC_SCALE: {v}", "LATVEC:", "1.000000000000000 0.000000000000000 0.000000000000000",
                      "0.000000000000000 1.000000000000000 0.000000000000000",
                      "0.000000000000000 0.000000000000000 1.000000000000000"]
        else:
            lines.append(f"{k}: {v}")
    lines += [bar, f"{'Cell':^75s}".rstrip(), bar, "Lattice vectors (Bohr):"]
    for a in range(3):
        row = [0.0, 0.0, 0.0]
        row[a] = L[a]
        lines.append(" ".join(f"{v:.15f}" for v in row))
    lines.append(f"Volume: {L[0] * L[1] * L[2]:.10E} (Bohr^3)")
    lines += [bar, f"{'Initialization':^75s}".rstrip(), bar,
              f"Mesh spacing                       :  {g.hx:.5g} (Bohr)",
              f"Total number of atom types         :  {len(case.type_names)}",
              f"Total number of atoms              :  {n_atom}",
              f"Total number of electrons          :  {case.params['nelec']:.10g}"]
    for t, (nm, cnt, f) in enumerate(zip(case.type_names, case.type_counts, case.psp_files)):
        lines += [f"Atom type {t + 1}                        :  {nm}",
                  f"Pseudopotential                    :  {f}",
                  f"Number of atoms of type {t + 1}          :  {cnt}"]
    lines += ["=" * 67, f"{'Self Consistent Field (SCF#1)':^67s}".rstrip(), "=" * 67,
              "Iteration     Free Energy (Ha/atom)   SCF Error        Timing (sec)"]
    for k, (e, err, ms) in enumerate(zip(result["iter_etot"], result["iter_err"], result["iter_ms"])):
        lines.append(f"{k + 1:<6d}       {e / n_atom:18.10E}        {err:.3E}        {ms * 1e-3:.3f}")
    lines += [f"Total number of SCF: {int(result['scf_iters']):<6d}", "=" * 68,
              f"{'Energy and force calculation':^68s}".rstrip(), "=" * 68]
    rows = [("Free energy per atom", result["Etot"] / n_atom, "(Ha/atom)"),
            ("Total free energy", result["Etot"], "(Ha)"),
            ("Band structure energy", result["Eband"], "(Ha)"),
            ("Exchange correlation energy", result["Exc"], "(Ha)"),
            ("Self and correction energy", case.params["Esc"], "(Ha)"),
            ("-Entropy*kb*T", result["Entropy"], "(Ha)"),
            ("Fermi level", result["Efermi"], "(Ha)")]
    for label, v, unit in rows:
        lines.append(f"{label:<35s}:{v:18.10E} {unit}")
    if "forces" in result:
        # printed as "RMS force": the mean of |F_J| (src/electronicGroundState.c:232-242)
        fn = np.linalg.norm(np.asarray(result["forces"]), axis=1)
        lines += [f"{'RMS force':<35s}:{fn.mean():18.10E} (Ha/Bohr)",
                  f"{'Maximum force':<35s}:{fn.max():18.10E} (Ha/Bohr)"]
    Path(path).write_text("\n".join(lines) + "\n")


def write_static(path, case: Case, result: dict) -> None:
    """The .static file: fractional atom positions per type, the total free
    energy and the atomic forces, in the reference's layout (no stress)."""
    g = case.grid
    L = np.array([g.hx * g.nx, g.hy * g.ny, g.hz * g.nz])
    bar = "*" * 75
    lines = [bar, f"{'Atom positions':^75s}".rstrip(), bar]
    for nm, t in zip(case.type_names, case.setup.types):
        lines.append(f"Fractional coordinates of {nm}:")
        for r in np.asarray(t.positions, float) / L:
            lines.append("".join(f"{v:18.10f}" for v in r))
    lines.append(f"Total free energy (Ha): {result['Etot']:.15E}")
    if "forces" in result:
        lines.append("Atomic forces (Ha/Bohr):")
        lines += ["".join(f"{v:18.10E}" for v in row) for row in np.asarray(result["forces"])]
    Path(path).write_text("\n".join(lines) + "\n")


def write_eigen(path, case: Case, result: dict) -> None:
    """The .eigen file: eigenvalues (Ha) and occupations (with the spin
    factor) of the Gamma point, in the reference's layout."""
    lam = np.asarray(result["lambda_"])
    occ = 2.0 * occupations(lam, result["Efermi"], case.params["beta"], case.params["smear_type"])
    lines = ["Final eigenvalues (Ha) and occupation numbers", "",
             "kred #1 = (0.000000,0.000000,0.000000)", "weight = 1.000000",
             "n        eigval                 occ"]
    lines += [f"{i + 1:<7d}{e:20.12E} {o:18.12f}" for i, (e, o) in enumerate(zip(lam, occ))]
    Path(path).write_text("\n".join(lines) + "\n")


def read_out_energies(path) -> dict:
    """The energy block and SCF table of a .out file (the reference's or
    write_out's) -> {label: value, 'scf': [(iter, E/atom, error)]}."""
    text = Path(path).read_text()
    out = {}
    for label in ("Free energy per atom", "Total free energy", "Band structure energy",
                  "Exchange correlation energy", "Self and correction energy", "-Entropy*kb*T",
                  "Fermi level"):
        m = re.search(re.escape(label) + r"\s*:\s*([-+0-9.Ee]+)", text)
        if m:
            out[label] = float(m.group(1))
    out["scf"] = [(int(a), float(b), float(c)) for a, b, c in
                  re.findall(r"^(\d+)\s+([-+0-9.E]+)\s+([-+0-9.E]+)\s+[0-9.]+\s*$", text, re.M)]
    return out


def run_case(case: Case, dtype: int = 0, poisson_net=None, ctx=None, forces=None) -> dict:
    """The device SCF of a loaded case (one context on cuda:0 unless given),
    followed by the atomic forces when forces is true (default: the case's
    PRINT_FORCES): result["forces"] [n_atom, 3] Ha/Bohr and their parts in
    result["forces_parts"]."""
    from . import gpu
    from .forces import scf_forces
    want = case.print_forces if forces is None else bool(forces)
    own = ctx is None
    if own:
        ctx = gpu.Context(case.grid)
    try:
        ctx.set_nonlocal(case.setup.nloc)
        r = ctx.scf_run(case.setup.b, case.setup.rho_at, dict(case.params, dtype=dtype),
                        rho_core=case.setup.rho_core, poisson_net=poisson_net, keep_psi=want)
        if want:
            f = scf_forces(ctx, case.setup, case.params, r, ref_cutoff=case.ref_cutoff)
            r["forces"] = f["total"]
            r["forces_parts"] = {k: v for k, v in f.items() if k != "total"}
            r.pop("psi").free()
        return r
    finally:
        if own:
            ctx.close()
