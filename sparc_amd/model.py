"""Silicon-like model systems for the SCF loop.

The ions are Gaussian pseudocharges of charge Z (b_J = -Z g_sigma(r - R_J)),
whose electrostatic self-energy Z^2 / (2 sqrt(pi) sigma) is removed through
Esc, plus Kleinman-Bylander projectors exp(-r^2 / 2 s^2) Y_lm (l <= 1) on every
atom.  Positions follow the diamond-cubic cells of the silicon workloads
(8, 64 or 216 atoms, a = 10.26 Bohr per cubic cell, seeded displacements).
The model exercises every step of SPARC's SCF loop on the device with inputs
that need no pseudopotential files.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .workloads import Grid, NonlocalSet, _periodic_delta, gaussian_projectors


def diamond_cubic(ncell: int, a: float = 10.26, disp: float = 0.02, seed: int = 0) -> np.ndarray:
    """Cartesian positions of an ncell^3 supercell of the 8-atom diamond cell."""
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    cells = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)])
    frac = (cells[:, None, :] + base[None, :, :]).reshape(-1, 3)
    pos = frac * a
    rng = np.random.default_rng(seed)
    return pos + rng.normal(0.0, disp, size=pos.shape)


def gaussian_density(grid: Grid, centers: np.ndarray, charges, sigma: float, cut: float = 6.0):
    """sum_J q_J (2 pi sigma^2)^(-3/2) exp(-|r - R_J|^2 / (2 sigma^2)), minimum image."""
    X, Y, Z = grid.coords()
    L = grid.lengths
    out = np.zeros(grid.nd)
    norm = (2 * np.pi * sigma * sigma) ** -1.5
    for c, q in zip(centers, np.broadcast_to(charges, (len(centers),))):
        dx = _periodic_delta(X - c[0], L[0], not grid.bcx)
        dy = _periodic_delta(Y - c[1], L[1], not grid.bcy)
        dz = _periodic_delta(Z - c[2], L[2], not grid.bcz)
        r2 = dx * dx + dy * dy + dz * dz
        msk = r2 < (cut * sigma) ** 2
        out[msk] += q * norm * np.exp(-r2[msk] / (2 * sigma * sigma))
    return out


@dataclass
class ModelSystem:
    grid: Grid
    positions: np.ndarray
    zval: float
    b: np.ndarray          # pseudocharge density, integrates to -nelec
    rho0: np.ndarray       # initial electron density, integrates to nelec
    Esc: float
    nloc: NonlocalSet
    nelec: float
    nstates: int
    meta: dict = field(default_factory=dict)

    @property
    def n_atom(self) -> int:
        return len(self.positions)


def merge_nonlocal(sets) -> NonlocalSet:
    """Concatenate projector sets (offsets shifted, alpha rows renumbered)."""
    sets = [q for q in sets if q is not None and q.n_influence > 0]
    po = co = ao = 0
    parts = {k: [] for k in ("ndc", "nproj", "pos_off", "chi_off", "alpha_off", "grid_pos", "chi",
                             "gamma")}
    for q in sets:
        parts["ndc"].append(q.ndc); parts["nproj"].append(q.nproj)
        parts["pos_off"].append(q.pos_off + po); parts["chi_off"].append(q.chi_off + co)
        parts["alpha_off"].append(q.alpha_off + ao)
        parts["grid_pos"].append(q.grid_pos); parts["chi"].append(q.chi)
        parts["gamma"].append(q.gamma)
        po += q.grid_pos.size; co += q.chi.size; ao += q.n_alpha
    cat = {k: np.concatenate(v) for k, v in parts.items()}
    return NonlocalSet(ndc=cat["ndc"].astype(np.int32), nproj=cat["nproj"].astype(np.int32),
                       pos_off=cat["pos_off"].astype(np.int64),
                       grid_pos=cat["grid_pos"].astype(np.int32),
                       chi_off=cat["chi_off"].astype(np.int64), chi=cat["chi"],
                       alpha_off=cat["alpha_off"].astype(np.int32), n_alpha=ao,
                       gamma=cat["gamma"], dV=sets[0].dV)


# Per-species model parameters: valence charge, pseudocharge width, initial
# density width, projector width and KB energies (l = 0, 1 share them).
SPECIES = {
    "Si": dict(zval=4.0, sigma_b=0.75, sigma_rho=1.2, proj_sigma=1.0, gammas=(2.0,)),
    "Al": dict(zval=3.0, sigma_b=0.85, sigma_rho=1.3, proj_sigma=1.1, gammas=(1.5,)),
}


def silicon_model(ncell: int = 1, n: int = 26, nstates: int | None = None, a: float = 10.26,
                  disp: float = 0.02, seed: int = 0, al_frac: float = 0.0, proj_rc: float = 3.0,
                  lmax: int = 1) -> ModelSystem:
    """Diamond-cubic ncell^3 silicon-like model on an n^3 periodic grid, with
    round(al_frac * n_atom) seeded Al substitutions."""
    L = a * ncell
    h = L / n
    g = Grid(n, n, n, h, h, h)
    pos = diamond_cubic(ncell, a, disp, seed) % L
    nat = len(pos)
    kind = np.array(["Si"] * nat, dtype=object)
    n_al = int(round(al_frac * nat))
    if n_al:
        kind[np.random.default_rng(seed + 1).choice(nat, n_al, replace=False)] = "Al"
    b = np.zeros(g.nd)
    rho0 = np.zeros(g.nd)
    esc, nelec, nls = 0.0, 0.0, []
    for sp, prm in SPECIES.items():
        sel = pos[kind == sp]
        if len(sel) == 0:
            continue
        z = prm["zval"]
        nelec += z * len(sel)
        b += gaussian_density(g, sel, -z, prm["sigma_b"])
        rho0 += gaussian_density(g, sel, z, prm["sigma_rho"])
        esc -= len(sel) * z * z / (2.0 * np.sqrt(np.pi) * prm["sigma_b"])
        nls.append(gaussian_projectors(g, sel, min(proj_rc, 0.49 * L),
                                       sigmas=(prm["proj_sigma"],) * len(prm["gammas"]),
                                       gammas=prm["gammas"], lmax=lmax))
    b *= -nelec / (b.sum() * g.dV)
    rho0 *= nelec / (rho0.sum() * g.dV)
    if nstates is None:
        nstates = int(int(nelec) // 2 * 1.2 + 5)  # SPARC's default Nstates (src/initialization.c:1716)
    return ModelSystem(g, pos, float(SPECIES["Si"]["zval"]), b, rho0, esc, merge_nonlocal(nls),
                       nelec, nstates,
                       meta={"ncell": ncell, "n": n, "a": a, "h": h, "n_al": n_al,
                             "species": list(kind)})


def scf_config(name: str) -> tuple[ModelSystem, dict]:
    """The SCF workloads by name: 'si8' (26^3, 24 states, Gaussian smearing
    0.2 eV), 'si64' (52^3, 160 states) and 'sial216' (78^3, 480 states, 10 %
    Al, 0.03 Bohr displacements, Fermi-Dirac smearing 0.01 Ha)."""
    if name == "si8":
        sysm = silicon_model(1, 26, 24)
        return sysm, default_params(sysm)
    if name == "si64":
        sysm = silicon_model(2, 52, 160)
        return sysm, default_params(sysm)
    if name == "sial216":
        sysm = silicon_model(3, 78, 480, disp=0.03, al_frac=0.1)
        p = default_params(sysm)
        p.update(smear_type=0, beta=100.0)
        return sysm, p
    raise ValueError(f"unknown SCF workload {name!r}")


def default_params(sysm: ModelSystem, precision: int = 0, tol_scf: float | None = None,
                   cheb_degree: int | None = None, max_scf: int = 100) -> dict:
    """SPARC's defaults for the SCF parameters (src/initialization.c:677-989,2233-2608)."""
    h = sysm.grid.hx
    if cheb_degree is None:
        p3, p2, p1, p0 = -700. / 3., 1240. / 3., -773. / 3., 1078. / 15.
        cheb_degree = 14 if h > 0.7 else int(round(((p3 * h + p2) * h + p1) * h + p0))
    if tol_scf is None:
        tol_scf = float(np.exp((np.log(1e-5) - 1.165) / 1.502))  # 1e-5 Ha/atom target
    return dict(nstates=sysm.nstates, cheb_degree=cheb_degree, rho_trigger=4, max_scf=max_scf,
                min_scf=2, tol_scf=tol_scf, tol_poisson=tol_scf * 0.01, maxit_poisson=3000,
                tol_lanczos=1e-2, beta=1.0 / (0.2 / 27.211386245988), smear_type=1,
                nelec=sysm.nelec, occfac=2.0, mix_history=7, pulay_freq=1, mix_beta=0.3,
                mix_omega=0.3, kerker_ktf=1.0, kerker_thresh=0.1, tol_precond=h * h * 1e-3,
                Esc=sysm.Esc, xc=0, dtype=precision, psi_seed=1234, lanczos_seed=4321)
