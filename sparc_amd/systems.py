"""SCF workloads built from psp8 pseudopotentials.

* ``sic8``: the reference's SiC_orthogonal_quick_relax test input (cubic
  8.2773683 Bohr, 4 Si + 4 C, FD_GRID 24^3 from MESH_SPACING 0.35, order 12,
  LDA_PW, Gaussian smearing 0.2 eV, TOL_SCF 1e-6, TOL_PSEUDOCHARGE 1e-9,
  24 states, degree 23) -- its reference output holds the energies.
* ``si8`` / ``si64`` / ``sial216``: diamond-cubic silicon, a = 10.26 Bohr per
  cubic cell, 1 / 2^3 / 3^3 cells on 26^3 / 52^3 / 78^3 grids (h = 0.3946),
  24 / 160 / 480 states, LDA_PZ, seeded displacements (0.02 / 0.02 /
  0.03 Bohr); sial216 has 10 % seeded Al substitutions and Fermi-Dirac
  smearing 0.01 Ha.  ``sial216_80``: the 216-atom cell on 80^3 (h = 0.3848),
  ``si8u``: the 8-atom cell on 28^3 (grids whose dimensions are multiples of
  4, on which the learned preconditioner runs without interpolation).
  Pseudopotentials: the ONCV psp8 files of SPARC's psps/ directory.
"""
from __future__ import annotations

import numpy as np

from .atoms import AtomType, setup_atoms
from .model import diamond_cubic
from .psp import read_psp8
from .workloads import Grid

SI = "14_Si_4_1.9_1.9_pbe_n_v1.0.psp8"
AL = "13_Al_3_1.9_1.9_pbe_n_v1.0.psp8"
C = "06_C_4_1.2_1.2_pbe_n_v1.0.psp8"
CONST_EH = 27.211386245988  # Hartree in eV (src/include/isddft.h)


def mesh2chebdegree(h: float) -> int:
    """Mesh2ChebDegree (src/eigenSolver.c:2347-2362): cubic fit in h."""
    p3, p2, p1, p0 = -700. / 3., 1240. / 3., -773. / 3., 1078. / 15.
    return 14 if h > 0.7 else int(round(((p3 * h + p2) * h + p1) * h + p0))


def tol_scf_for(de_per_atom: float) -> float:
    """Default TOL_SCF for a target energy accuracy (src/initialization.c:2551-2555)."""
    return float(np.exp((np.log(de_per_atom) - 1.165) / 1.502))


def scf_params(setup, nstates: int, h: float, tol_scf: float, xc: int, smear_type: int = 1,
               beta: float | None = None, cheb_degree: int | None = None, dtype: int = 0,
               tol_poisson: float | None = None) -> dict:
    if beta is None:
        beta = CONST_EH / 0.2 if smear_type == 1 else CONST_EH / 0.1
    return dict(nstates=nstates, cheb_degree=cheb_degree or mesh2chebdegree(h), rho_trigger=4,
                max_scf=100, min_scf=2, tol_scf=tol_scf,
                tol_poisson=tol_poisson if tol_poisson is not None else tol_scf * 0.01,
                maxit_poisson=3000, tol_lanczos=1e-2, beta=beta, smear_type=smear_type,
                nelec=setup.nelec, occfac=2.0, mix_history=7, pulay_freq=1, mix_beta=0.3,
                mix_omega=0.3, kerker_ktf=1.0, kerker_thresh=0.1, tol_precond=h * h * 1e-3,
                Esc=setup.Esc, xc=xc, dtype=dtype, psi_seed=1234, lanczos_seed=4321)


def sic8():
    """(setup, params) of the SiC_orthogonal_quick_relax reference test."""
    L = 8.2773683
    n = 24
    h = L / n
    si = np.array([[0.26, 0.26, 0.75], [0.25, 0.75, 0.25], [0.75, 0.25, 0.25],
                   [0.75, 0.75, 0.75]]) * L
    c = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.51, 0.51, 0]]) * L
    g = Grid(n, n, n, h, h, h)
    types = [AtomType(read_psp8(SI), si, "Si"), AtomType(read_psp8(C), c, "C")]
    st = setup_atoms(g, types, tol_pseudocharge=1e-9, ref_cutoff=0.5)
    p = scf_params(st, 24, h, 1e-6, xc=1, smear_type=1, beta=1.0 / 0.007349864435)
    return st, p


def silicon(ncell: int, n: int, nstates: int, disp: float = 0.02, seed: int = 0,
            al_frac: float = 0.0, fermi_dirac: float | None = None):
    """(setup, params) of a diamond-cubic Si (optionally Si/Al) supercell."""
    a = 10.26
    L = a * ncell
    h = L / n
    g = Grid(n, n, n, h, h, h)
    pos = diamond_cubic(ncell, a, disp, seed) % L
    nat = len(pos)
    is_al = np.zeros(nat, bool)
    n_al = int(round(al_frac * nat))
    if n_al:
        is_al[np.random.default_rng(seed + 1).choice(nat, n_al, replace=False)] = True
    types = [AtomType(read_psp8(SI), pos[~is_al], "Si")]
    if n_al:
        types.append(AtomType(read_psp8(AL), pos[is_al], "Al"))
    tol = tol_scf_for(1e-5)
    st = setup_atoms(g, types, tol_pseudocharge=tol * 0.001, ref_cutoff=0.5)
    st.meta.update(ncell=ncell, n=n, n_al=n_al)
    if fermi_dirac is not None:
        p = scf_params(st, nstates, h, tol, xc=0, smear_type=0, beta=1.0 / fermi_dirac)
    else:
        p = scf_params(st, nstates, h, tol, xc=0, smear_type=1)
    return st, p


def scf_system(name: str):
    if name == "sic8":
        return sic8()
    if name == "si8":
        return silicon(1, 26, 24)
    if name == "si64":
        return silicon(2, 52, 160)
    if name == "si8u":
        return silicon(1, 28, 24)
    if name == "sial216":
        return silicon(3, 78, 480, disp=0.03, al_frac=0.1, fermi_dirac=0.01)
    if name == "sial216_80":
        return silicon(3, 80, 480, disp=0.03, al_frac=0.1, fermi_dirac=0.01)
    raise ValueError(f"unknown SCF system {name!r}")
